"""GPU timings in the reference optimizer's cache format (exp/preCompute<crypto>.abs).

The reference's PIROptimizer fills that cache with, per cryptosystem parameter string,
the CPU time to absorb-and-reply one plaintext and the time to pre-compute (import)
one plaintext (pir/replyGenerator/PIROptimizer.cpp:120-210); its client-side optimizer
reads them to choose d, alpha and the parameters.  The lines written here have the same
form -- "<param> <abs1plaintext_s> <precompute1plaintext_s>", 5 significant digits
(PIROptimizer.cpp:20,142-145) -- with the times measured on the GPU, so that optimizer
picks parameters for this server.

Measurement follows the reference's definitions:
  * a plaintext is n*m*8 bytes (one polynomial of 64-bit words;
    PIRReplyGeneratorNFL_internal.cpp:62-73);
  * d = 1, alpha = 1, n0 = 4 elements of plaintext_nbr plaintexts each, A from n0 = 4;
  * abs: reply time, plaintext_nbr from 4 doubling while the time is under 0.5 s and
    plaintext_nbr * 4 * ciphertext_bits < 2^30, then time / (4 * plaintext_nbr)
    (PIROptimizer.cpp:155-180);
  * precompute: import time, the same loop from 128 (PIROptimizer.cpp:182-207).
The set of parameter strings is the caller's: the reference enumerates it from its
security-estimate table (crypto/NFLLWE.cpp:554-588), which is not restated here; pass
the names, or read them from an existing cache file (--like).

    python -m xpir_amd.optim OUT.abs LWE:80:1024:60 LWE:80:2048:120 ...
    python -m xpir_amd.optim OUT.abs --like exp/preComputeLWE.abs
"""
from __future__ import annotations

import argparse
import sys
import time

N0 = 4
LIMIT_S = 0.5
SIZE_CAP_BITS = 1 << 30


def doubling_loop(simulate, start: int, ct_bits: int, stop_on_zero: bool = False) -> float:
    """The reference's search: time `simulate(plaintext_nbr)` for plaintext_nbr = start,
    2*start, ... while under LIMIT_S seconds and under the size cap; returns seconds per
    plaintext of the last run."""
    p = start
    while True:
        result = simulate(p)
        p *= 2
        if not (result < LIMIT_S and (not stop_on_zero or result != 0.0) and p * N0 * ct_bits < SIZE_CAP_BITS):
            break
    p //= 2
    return result / (N0 * p)


def format_line(param: str, abs_s: float, pre_s: float) -> str:
    return f"{param} {abs_s:.5g} {pre_s:.5g}"


def params_from_file(path: str):
    """Parameter names of an existing cache file (first field of each line)."""
    out = []
    with open(path) as f:
        for line in f:
            fields = line.split(" ")
            if fields and fields[0].strip():
                out.append(fields[0].strip())
    return out


def measure(ctx, param: str, reps: int = 3):
    """(abs seconds per plaintext, precompute seconds per plaintext) on ctx's GPU."""
    from . import capi
    prm = capi.params(param, N0)
    plain_bytes = prm.n * prm.m * 8
    ct_bits = prm.ct_bytes * 8
    sk = capi.keygen(prm, 1)
    skd = ctx.alloc(sk.nbytes).upload(sk)

    def import_db(p):
        elem = p * plain_bytes
        raw = ctx.alloc(N0 * elem + 16)
        ctx.random_bytes_dev(7, 0, raw, N0 * elem)
        ctx.sync()
        t0 = time.perf_counter()
        db = ctx.import_db_dev(prm, raw, N0, elem, [N0])
        ctx.sync()
        dt = time.perf_counter() - t0
        raw.free()
        return db, dt

    def reply_time(p):
        db, _ = import_db(p)
        qd = ctx.alloc(db.query_words * 8)
        ctx.query_gen_dev(prm, skd, [N0], 0, 3, qd)
        rd = ctx.alloc(db.reply_words * 8)
        ctx.reply_dev(db, qd, rd)  # first call sizes the context buffers
        ctx.sync()
        best = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            ctx.reply_dev(db, qd, rd)
            ctx.sync()
            best = min(best, time.perf_counter() - t0)
        for b in (qd, rd):
            b.free()
        db.free()
        return best

    def import_time(p):
        best = 1e30
        for _ in range(reps):
            db, dt = import_db(p)
            db.free()
            best = min(best, dt)
        return best

    try:
        abs_s = doubling_loop(reply_time, 4, ct_bits)
        pre_s = doubling_loop(import_time, 128, ct_bits, stop_on_zero=True)
    finally:
        skd.free()
    return abs_s, pre_s


def write_abs(ctx, params, path: str):
    lines = [format_line(p, *measure(ctx, p)) for p in params]
    with open(path, "w") as f:
        f.write("".join(line + "\n" for line in lines))
    return lines


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("out")
    ap.add_argument("params", nargs="*")
    ap.add_argument("--like", help="take the parameter names from this cache file")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    params = list(a.params) + (params_from_file(a.like) if a.like else [])
    if not params:
        ap.error("no parameter strings")
    from . import capi
    ctx = capi.Context(a.device)
    try:
        for line in write_abs(ctx, params, a.out):
            print(line)
    finally:
        ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
