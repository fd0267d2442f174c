"""Write the golden fixtures of tests/golden/ from the REFERENCE's own arithmetic.

Server-side words (import = bitsplitter + nttAndPowPhi, reply = mulandaddShoup chains,
recursion = deserializeDataNFL over the level's reply buffers) come from
oracle/_ref/libxpir_ref.so, the reference's NFLlib compiled from its sources by
oracle/Makefile.  The reference's query generator (NFLLWE::enc) needs Boost, which the
image lacks, so queries are produced by the oracle's encryption; they are inputs here,
not checked outputs.

    python tests/golden/make_golden.py            # C1, C3-scaled, absorb and alpha vectors
    python tests/golden/make_golden.py --c2       # also the 65536 x 4 KiB single query
    python tests/golden/make_golden.py --full     # also the recursive 2^20 x 1 KiB grid and
                                                  # the eight batched databases, row-streamed

The full-size grids do not fit in host memory as NTT databases (96 GiB for the recursive
configuration), so their level 0 is row-streamed: the reference's deserializeDataNFL and
mulandaddShoup run one level-0 group (n0 elements) at a time on 8 threads
(RefNFL.level0_stream); the level-0 replies then go through the reference's recursion
import and the last level as for the small grids.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle, RefNFL, build  # noqa: E402
from xpir_amd.workloads import (C1_PLAIN, C2_SINGLE, C3_RECURSIVE, C3_SCALED, batch_databases,  # noqa: E402
                                batch_golden_query, serving_d1_views)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def ref_generate(cfg, threads=8):
    """The reference's arithmetic for a whole d-level reply on a seeded workload."""
    desc, dims = cfg["crypto"], cfg["dims"]
    orc = Oracle(desc, dims[0])
    ref = RefNFL(orc.n, orc.m * 60)
    n, L = cfg["n_elems"], cfg["elem_bytes"]
    data = orc.rand_bytes(cfg["db_seed"], 0, n * L)
    P0 = orc.polys(L)
    units = int(np.prod(dims))
    words = np.zeros(units * P0 * orc.mn, dtype=np.uint64)
    for e in range(n):
        words[e * P0 * orc.mn:(e + 1) * P0 * orc.mn] = ref.import_buffers(data[e * L:(e + 1) * L], 1, orc.A)
    sk = orc.keygen(cfg["key_seed"])
    q = orc.make_query(sk, dims, cfg["index"], cfg["noise_seed"])
    level_words, P = words, P0
    samples = {"ntt_unit0_head": [int(x) for x in words[:8]],
               "ntt_last_head": [int(x) for x in words[(n - 1) * P0 * orc.mn:(n - 1) * P0 * orc.mn + 8]],
               "ntt_sha256": sha(words)}
    for lvl, nl in enumerate(dims):
        groups = int(np.prod(dims[lvl + 1:], dtype=np.int64)) if lvl + 1 < len(dims) else 1
        rep = ref.reply_level(level_words, q[lvl], nl, groups, P, threads)
        if lvl + 1 < len(dims):
            ct_bytes = P * 2 * orc.mn * 8
            P2 = orc.next_polys(P)
            nxt = np.zeros(groups * P2 * orc.mn, dtype=np.uint64)
            rb = rep.view(np.uint8)
            for g in range(groups):
                nxt[g * P2 * orc.mn:(g + 1) * P2 * orc.mn] = ref.import_buffers(
                    rb[g * ct_bytes:(g + 1) * ct_bytes], P, orc.A)
            samples[f"level{lvl}_reply_sha256"] = sha(rep)
            level_words, P = nxt, P2
        else:
            final = rep
    entry = dict(cfg, A=orc.A, P0=P0, Pd=P, reply_sha256=sha(final),
                 reply_head=[int(x) for x in final[:16]], reply_tail=[int(x) for x in final[-16:]],
                 noncanonical_words=int(np.sum(final >= np.tile(np.repeat(np.array(orc.p, dtype=np.uint64), orc.n),
                                                                2 * P))),
                 **samples)
    return entry, final


def ref_generate_streamed(cfg, threads=8):
    """The reference's arithmetic for a d = 2 reply on a seeded workload too large to hold:
    level 0 row-streamed group by group, then the recursion and level 1 whole."""
    import time
    desc, dims = cfg["crypto"], cfg["dims"]
    assert len(dims) == 2
    orc = Oracle(desc, dims[0])
    ref = RefNFL(orc.n, orc.m * 60)
    n, L = cfg["n_elems"], cfg["elem_bytes"]
    P0 = orc.polys(L)
    sk = orc.keygen(cfg["key_seed"])
    t0 = time.perf_counter()
    q = orc.make_query(sk, dims, cfg["index"], cfg["noise_seed"])
    t1 = time.perf_counter()
    groups = dims[1]
    rep0 = ref.level0_stream(orc.A, cfg["db_seed"], n, L, dims[0], 0, groups, P0, q[0], threads)
    t2 = time.perf_counter()
    first = orc.rand_bytes(cfg["db_seed"], 0, L)
    last = orc.rand_bytes(cfg["db_seed"], (n - 1) * L, L)
    w_first, w_last = ref.import_buffers(first, 1, orc.A), ref.import_buffers(last, 1, orc.A)
    ct_bytes = P0 * 2 * orc.mn * 8
    P1 = orc.next_polys(P0)
    nxt = np.zeros(groups * P1 * orc.mn, dtype=np.uint64)
    rb = rep0.view(np.uint8)
    for g in range(groups):
        nxt[g * P1 * orc.mn:(g + 1) * P1 * orc.mn] = ref.import_buffers(rb[g * ct_bytes:(g + 1) * ct_bytes], P0, orc.A)
    final = ref.reply_level(nxt, q[1], dims[1], 1, P1, threads)
    t3 = time.perf_counter()
    g_words = P0 * 2 * orc.mn
    entry = dict(cfg, A=orc.A, P0=P0, Pd=P1, reply_sha256=sha(final),
                 reply_head=[int(x) for x in final[:16]], reply_tail=[int(x) for x in final[-16:]],
                 ntt_unit0_head=[int(x) for x in w_first[:8]], ntt_last_head=[int(x) for x in w_last[:8]],
                 level0_reply_sha256=sha(rep0),
                 level0_group0_head=[int(x) for x in rep0[:8]],
                 level0_last_group_tail=[int(x) for x in rep0[(groups - 1) * g_words:][-8:]],
                 level1_ntt_sha256=sha(nxt),
                 noncanonical_words=int(np.sum(final >= np.tile(np.repeat(np.array(orc.p, dtype=np.uint64), orc.n),
                                                                2 * P1))))
    print(f"  {cfg['name']}: query {t1 - t0:.1f} s, level 0 streamed in {t2 - t1:.1f} s, level 1 {t3 - t2:.1f} s "
          f"({threads} threads), reply sha {entry['reply_sha256'][:16]}", flush=True)
    return entry


def batch_goldens(threads=8):
    out = []
    for i, spec in enumerate(batch_databases()):
        cfg = dict(spec, name=f"batched-db{i}-{spec['n_elems']}x{spec['elem_bytes']}", **batch_golden_query(i, spec))
        out.append(ref_generate_streamed(cfg, threads))
    return out


def absorb_vectors():
    """Bitsplitter vectors: A = 15 / 42 / 74, single buffers and buffers whose chunks
    straddle buffer boundaries (the recursion layout)."""
    out = []
    for desc, n0, nbufs, per in [("LWE:80:1024:60", 1000, 1, 10240), ("LWE:80:1024:60", 1000, 3, 16384),
                                 ("LWE:80:2048:120", 65536, 1, 4096), ("LWE:80:2048:120", 65536, 2, 65536),
                                 ("LWE:100:4096:180", 1024, 1, 1024), ("LWE:100:4096:180", 1024, 3, 196608),
                                 ("LWE:80:2048:120", 7, 1, 4101)]:
        orc = Oracle(desc, n0)
        ref = RefNFL(orc.n, orc.m * 60)
        data = orc.rand_bytes(99, 0, nbufs * per)
        words = ref.import_buffers(data, nbufs, orc.A)
        out.append(dict(crypto=desc, n0=n0, nbufs=nbufs, bytes_per_buf=per, seed=99, A=orc.A,
                        polys=int(words.size // orc.mn), sha256=sha(words), head=[int(x) for x in words[:8]]))
    return out


ALPHA_CASES = [("LWE:80:1024:60", 4, 10, 10240, 3), ("LWE:80:1024:60", 4, 11, 10240, 3),
               ("LWE:80:2048:120", 4, 10, 4096, 3), ("LWE:80:2048:120", 1, 2, 4096, 3),
               ("LWE:80:2048:120", 3, 9, 4096, 2)]


def alpha_vectors():
    """Imports with alpha aggregation where n_elems is not a multiple of alpha: the units
    come from the reference's DBHandler::readAggregatedStream over a DBVectorProcessor in
    importDataNFL's loop (one reused buffer), then the reference's deserializeDataNFL."""
    out = []
    for desc, n0, n, L, alpha in ALPHA_CASES:
        orc = Oracle(desc, n0)
        ref = RefNFL(orc.n, orc.m * 60)
        data = orc.rand_bytes(77, 0, n * L)
        units = RefNFL.aggregate([data[e * L:(e + 1) * L].tobytes() for e in range(n)], alpha)
        words = np.concatenate([ref.import_buffers(u, 1, orc.A) for u in units])
        out.append(dict(crypto=desc, n0=n0, n_elems=n, elem_bytes=L, alpha=alpha, seed=77, A=orc.A,
                        units=int(units.shape[0]), polys=int(words.size // orc.mn // units.shape[0]),
                        units_sha256=sha(units), ntt_sha256=sha(words),
                        last_unit_head=[int(x) for x in words[(units.shape[0] - 1) * (words.size // units.shape[0]):][:8]]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--c2", action="store_true")
    ap.add_argument("--full", action="store_true")
    ap.add_argument("--only-full", action="store_true", help="write only the full-size entries")
    args = ap.parse_args()
    build()
    assert RefNFL.available(), "oracle/_ref/libxpir_ref.so missing (needs the reference tree)"
    path = os.path.join(HERE, "golden.json")
    gold = json.load(open(path)) if os.path.exists(path) else {}
    if not args.only_full:
        e, rep = ref_generate(C1_PLAIN)
        gold["c1_plain"] = e
        np.save(os.path.join(HERE, "c1_plain_reply.npy"), rep)
        gold["c3_scaled"], _ = ref_generate(C3_SCALED)
        gold["absorb"] = absorb_vectors()
        gold["alpha"] = alpha_vectors()
        if args.c2:
            gold["c2_single"], _ = ref_generate(C2_SINGLE)
    if args.full or args.only_full:
        gold["c3_recursive"] = ref_generate_streamed(C3_RECURSIVE)
        json.dump(gold, open(path, "w"), indent=1)
        gold["c4_batched"] = batch_goldens()
        gold["c5_d1_views"] = [ref_generate(v)[0] for v in serving_d1_views()]
    json.dump(gold, open(path, "w"), indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
