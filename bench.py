#!/usr/bin/env python3
"""bench.py -- NFLLWE PIR reply throughput on one MI355X.

Workload (default): single query, LWE:80:2048:120 (n=2048, two 60-bit moduli), d=1,
alpha=1, 65536 elements x 4 KiB of seeded random bytes (256 MiB), one seeded query.
The database is imported (absorb + NTT) once and stays resident in HBM; the query
(65536 ciphertexts, 4 GiB) is generated on the GPU and resident before timing starts.

A step is one whole reply: xpg_reply_dev (query in HBM -> reply in HBM).
value = raw DB bytes / reply time (decimal GB/s), the reference's "DB processing
throughput" (PIRReplyGeneratorNFL_internal.cpp:286-288) in bytes instead of bits.

Also reported: p50/p99 reply latency, import GB/s, bit-exactness against the golden
reply bytes (tests/golden/golden.json, written from the reference's own arithmetic),
the MAC kernel's roofline (HIP events around every MAC launch on the context stream)
and the reference's CPU reply on this host (oracle/_ref, same inputs, same run).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(db_words: np.ndarray, query: np.ndarray, cfg, prm, gpu_reply: np.ndarray, raw_bytes: int):
    """The reference's MAC (its own NFLlib::mulandaddShoup, driven in generateReply's
    order) on this host's cores over the same NTT database and query."""
    try:
        from oracle.oracle import Oracle, RefNFL
    except Exception as e:  # pragma: no cover
        return {"error": f"oracle unavailable: {e}"}
    n_l = cfg["dims"][0]
    P = db_words.size // (n_l * prm.m * prm.n)
    if RefNFL.available():
        ref = RefNFL(prm.n, prm.m * 60)
        # the reference parallelises only over the P polys of an element (OpenMP,
        # PIRReplyGeneratorNFL_internal.cpp:531-534): P threads at most
        threads = max(1, min(P, os.cpu_count() or 1))
        rep, secs = ref.reply_level_timed(db_words, query, n_l, 1, P, threads)
        kind = "reference"
    else:
        orc = Oracle(cfg["crypto"], n_l)
        threads = 1
        t0 = time.perf_counter()
        rep = orc.reply_level(db_words, query, n_l, 1, P, threads)
        secs = time.perf_counter() - t0
        kind = "port"
    return {"value": raw_bytes / secs / 1e9, "unit": "GB/s", "cores": threads, "kind": kind,
            "seconds": secs, "matches_gpu_reply": bool(np.array_equal(rep, gpu_reply)),
            "sample": f"the full workload, one reply ({n_l} elements x {cfg['elem_bytes']} B), MAC loops "
                      f"only (the span of the reference's reply timer); {threads} thread(s) because the "
                      f"reference parallelises over the {P} poly(s) of an element"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="single", choices=["single", "plain"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import xpir_amd
    from xpir_amd.workloads import C1_PLAIN, C2_SINGLE

    cfg = C2_SINGLE if args.workload == "single" else C1_PLAIN
    gold_path = os.path.join(ROOT, "tests", "golden", "golden.json")
    gold = json.load(open(gold_path)) if os.path.exists(gold_path) else {}
    gold = gold.get("c2_single" if args.workload == "single" else "c1_plain")

    ctx = xpir_amd.Context(0)
    prm = xpir_amd.params(cfg["crypto"], cfg["dims"][0])
    n, L, dims = cfg["n_elems"], cfg["elem_bytes"], cfg["dims"]
    raw = n * L

    log(f"workload {cfg['name']}: generating {raw / 2**20:.0f} MiB of DB bytes on the GPU")
    dbb = ctx.alloc(raw + 16)
    ctx.random_bytes_dev(cfg["db_seed"], 0, dbb, raw)
    ctx.sync()
    t0 = time.perf_counter()
    db = ctx.import_db_dev(prm, dbb, n, L, dims)
    ctx.sync()
    import_s = time.perf_counter() - t0
    dbb.free()
    log(f"import (absorb + NTT) {import_s * 1e3:.1f} ms -> {raw / import_s / 1e9:.2f} GB/s raw")

    sk = xpir_amd.keygen(prm, cfg["key_seed"])
    skd = ctx.alloc(sk.nbytes).upload(sk)
    qwords = db.query_words
    qd = ctx.alloc(qwords * 8)
    ctx.query_gen_dev(prm, skd, dims, cfg["index"], cfg["noise_seed"], qd)
    rd = ctx.alloc(db.reply_words * 8)
    ctx.sync()
    log(f"query {qwords * 8 / 2**30:.2f} GiB generated on the GPU; warmup {args.warmup}")

    for _ in range(args.warmup):
        ctx.reply_dev(db, qd, rd)
    ctx.sync()
    reply = rd.download()
    digest = hashlib.sha256(reply.tobytes()).hexdigest()
    bit_exact = (gold is not None and digest == gold["reply_sha256"])
    log(f"reply sha256 {digest[:16]}... golden match: {bit_exact}")

    ctx.timing(True)
    ctx.timing_read(reset=True)
    lat = []
    ctx.sync()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        ctx.reply_dev(db, qd, rd)
        ctx.sync()
        lat.append(time.perf_counter() - t0)
    ctx.sync()
    total = time.perf_counter() - t_start
    mac_ms, launches = ctx.timing_read(reset=True)
    ctx.timing(False)
    ms_per_step = total / args.steps * 1e3
    value = raw / (total / args.steps) / 1e9
    log(f"{args.steps} replies: {ms_per_step:.3f} ms/reply -> {value:.2f} GB/s")

    mn = prm.m * prm.n
    P = db.polys
    ntt_bytes = n * P * mn * 8
    q_bytes = qwords * 8
    reply_bytes = db.reply_words * 8
    alg_bytes = ntt_bytes + q_bytes + reply_bytes
    # the MAC of one reply (every MAC launch of the reply, timed with HIP events)
    mac_reply_s = mac_ms / args.steps / 1e3
    achieved = alg_bytes / mac_reply_s / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "kernel": "reply MAC", "kernel_ms": mac_ms / args.steps,
                "launches_per_reply": launches / args.steps, "alg_bytes": alg_bytes}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline: the reference's MAC on this host over the same inputs")
        words = db.export_words()
        query = qd.download()
        cpu = cpu_baseline(words, query, cfg, prm, reply, raw)
        log(f"cpu baseline {cpu.get('value', 0):.3f} GB/s ({cpu.get('seconds', 0):.2f} s), "
            f"matches GPU reply: {cpu.get('matches_gpu_reply')}")

    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "DB GB/s per reply, batched vs plain, 1 MI355X, bit-exact; p50/p99 reply ms",
        "value": value, "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic (seeded SplitMix64 bytes)",
        "config": {"workload": f"{cfg['name']} single query d=1 alpha=1", "crypto": cfg["crypto"],
                   "n_elems": n, "elem_bytes": L, "dims": dims, "raw_db_bytes": raw,
                   "query_bytes": q_bytes, "ntt_db_bytes": ntt_bytes},
        "bit_exact": bool(bit_exact),
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "import_gbs": raw / import_s / 1e9,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    for b in (skd, qd, rd):
        b.free()
    db.free()
    ctx.close()


if __name__ == "__main__":
    main()
