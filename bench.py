#!/usr/bin/env python3
"""bench.py -- NFLLWE PIR reply throughput on one MI355X.

Default workload (the metric, "DB GB/s per reply, batched vs plain; p50/p99 reply ms"):
the batched uneven configuration -- eight resident LWE:80:2048:120 databases of
10 000 .. 1 000 000 elements of 256 B .. 16 KiB (seeded bytes), d = 2 grids, and the
seeded number of queries for each (334 in all), generated on the GPU and resident in HBM
before timing starts.

A step is one batched pass: every query answered by one xpg_reply_batch_dev call (query
in HBM -> reply in HBM).  value = raw DB bytes of all replies / pass time (decimal GB/s),
the reference's "DB processing throughput" (PIRReplyGeneratorNFL_internal.cpp:286-288)
in bytes.  Beside it: the plain path (each query alone through xpg_reply_dev, one after
another) on the same queries, and p50/p99 reply latency from a seeded Poisson stream
through the continuous-batching server over the same databases (plus d = 1 views of the
two smallest).  Bit-exactness: each database's golden query against the reference's
reply bytes (tests/golden/golden.json), every batched reply against the plain reply.
roofline: the batched level-0 kernel (mac_fused_kernel: one persistent launch over all
eight databases; --batch-fuse 0 times one launch per database chunk), integer-issue bound: exact
chain steps/s (HIP events around its launches in the timed passes) x VALU lane
instructions per step (rocprofv3 SQ_INSTS_VALU, profiles/) against 39.3 T lane ops/s.
cpu_baseline: the reference's own arithmetic (oracle/_ref) on this host on a bounded
sample of the same replies.  The single-query configuration (65536 x 4 KiB, HBM-bound)
is reported under "single_query".
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import shutil
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


def cpu_all_cores(db_words: np.ndarray, query: np.ndarray, cfg, prm, gpu_reply: np.ndarray, raw_bytes: int):
    """The clean-room restatement (oracle/xpir_oracle.c) with every host core busy: threads
    split over slot blocks, each chain still in element order, so the words are the
    reference's.  Timed: the MAC loops, the span of the reference's reply timer."""
    from oracle.oracle import Oracle
    n_l = cfg["dims"][0]
    P = db_words.size // (n_l * prm.m * prm.n)
    # the host share of one GPU: OMP_NUM_THREADS when set (16 per GPU), else every core
    threads = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
                         or (os.cpu_count() or 1)))
    orc = Oracle(cfg["crypto"], n_l)
    rep, secs = orc.reply_level_chains_timed(db_words, query, n_l, 1, P, threads)
    return {"value": raw_bytes / secs / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
            "seconds": secs, "matches_gpu_reply": bool(np.array_equal(rep, gpu_reply)),
            "sample": f"the full workload, one reply ({n_l} elements x {cfg['elem_bytes']} B), MAC loops "
                      f"only, {threads} threads over slot blocks"}


def measured_traffic(workload_name):
    """HBM bytes per reply of the MAC kernels, from the rocprofv3 counter runs committed
    under profiles/ (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the factor the
    same counter shows on a plain copy of known size); None when no summary is present."""
    path = os.path.join(ROOT, "profiles", "mac_traffic.json")
    try:
        rec = json.load(open(path))
    except (OSError, ValueError):
        return None
    if rec.get("workload") != workload_name:
        return None
    return rec.get("traffic_bytes_per_launch")


def bench_copy(args):
    """Counter calibration: a plain device copy of 4 GiB (reads 4 GiB, writes 4 GiB),
    repeated; profile it with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to get the bytes
    each counter unit stands for on this GPU."""
    import torch
    n = 4 << 30
    src = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    src.fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps)):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(1, args.steps)
    print(json.dumps({"workload": "copy", "bytes_read": n, "bytes_written": n, "ms_per_copy": dt * 1e3,
                      "gbs_read_plus_write": 2 * n / dt / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default 3 for metric, else 20)")
    ap.add_argument("--workload", default="metric",
                    choices=["metric", "batched-pass", "single", "plain", "recursive", "batched", "serving",
                             "session", "copy", "dbfile"])
    ap.add_argument("--scale", type=float, default=1.0, help="element-count scale of the batched/serving databases")
    ap.add_argument("--queries", type=int, default=None,
                    help="serving: queries in the Poisson stream (default 2000; metric: 3000)")
    ap.add_argument("--rate", type=float, default=None, help="serving: arrivals per second (default 300)")
    ap.add_argument("--no-single", action="store_true", help="metric: skip the single-query sub-report")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the MAC autotuner (fixed kernel rules); it runs before any timed region")
    ap.add_argument("--tune-file", help="install this saved tuning table instead of tuning when it exists; "
                                        "else tune and save the table there")
    ap.add_argument("--workers", type=int, default=2, help="serving: worker contexts (HIP streams) of the server")
    ap.add_argument("--heavy-steps", type=int, default=0,
                    help="serving: exact chain steps above which a batch is heavy (one heavy batch at a time; "
                         "0 = the server's default)")
    ap.add_argument("--batch-fuse", type=int, default=1, choices=(0, 1),
                    help="batched passes: level 0 of every database in fused persistent launches (1) or one "
                         "launch per database chunk (0)")
    ap.add_argument("--db-layout", type=int, default=1, choices=(0, 1),
                    help="batched/serving databases: level 0 on the matrix cores (1, the matrix layout) or "
                         "the row-major words with the VALU kernels (0)")
    ap.add_argument("--host-io", action="store_true",
                    help="single/plain: also time replies from a host query to a host reply (PCIe included)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: the engine has no multi-GPU path; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import xpir_amd
    from xpir_amd.workloads import C1_PLAIN, C2_SINGLE

    if args.warmup is None:
        args.warmup = 3 if args.workload == "metric" else 20
    if args.queries is None:
        args.queries = 3000 if args.workload == "metric" else 2000
    if args.rate is None:
        args.rate = 300.0
    if args.workload == "metric":
        return bench_metric(args)
    if args.workload == "batched-pass":
        return bench_batched_pass(args)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "recursive":
        return bench_recursive(args)
    if args.workload == "session":
        return bench_session(args)
    if args.workload == "dbfile":
        return bench_dbfile(args)

    cfg = C2_SINGLE if args.workload == "single" else C1_PLAIN
    gold = load_gold().get("c2_single" if args.workload == "single" else "c1_plain")
    ctx = xpir_amd.Context(0)
    out = run_single(ctx, args, cfg, gold)
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_dbfile(args):
    """The NTT database through a file against importing it again: import from raw bytes
    in HBM and from host memory, xpg_db_save (device -> file) and xpg_db_load (file ->
    device, the file just written so in the page cache, plus the device checksum), on the
    single-query database (65536 x 4 KiB, 2 GiB of NTT words) and the 100000 x 16 KiB
    batched database (6.1 GiB).  Every loaded database's words equal the imported ones'."""
    import tempfile

    import xpir_amd
    from xpir_amd.workloads import C2_SINGLE, batch_databases
    ctx = xpir_amd.Context(0)
    tmpdir = tempfile.mkdtemp(prefix="xpg_dbfile_")
    rows = []
    try:
        for spec in (C2_SINGLE, batch_databases()[3]):
            prm = xpir_amd.params(spec["crypto"], spec["dims"][0])
            n, L, dims = spec["n_elems"], spec["elem_bytes"], spec["dims"]
            raw = ctx.alloc(n * L + 16)
            ctx.random_bytes_dev(spec["db_seed"], 0, raw, n * L)
            db = ctx.import_db_dev(prm, raw, n, L, dims)  # builds the tables
            db.free()
            ctx.sync()
            t0 = time.perf_counter()
            db = ctx.import_db_dev(prm, raw, n, L, dims)
            ctx.sync()
            t_dev = time.perf_counter() - t0
            host = raw.download(np.uint8)[: n * L]
            raw.free()
            t0 = time.perf_counter()
            db_h = ctx.import_db(prm, host, n, L, dims)
            ctx.sync()
            t_host = time.perf_counter() - t0
            db_h.free()
            del host
            path = os.path.join(tmpdir, "db.ntt")
            digest = hashlib.sha256(f"{spec}".encode()).digest()
            t0 = time.perf_counter()
            ctx.save_db(db, path, digest)
            t_save = time.perf_counter() - t0
            t0 = time.perf_counter()
            db2 = ctx.load_db(prm, path, digest)
            t_load = time.perf_counter() - t0
            ntt = os.path.getsize(path) - 512
            same = bool(np.array_equal(db2.export_words(), db.export_words()))
            db2.free()
            db.free()
            os.remove(path)
            row = {"db": f"{n} x {L} B {spec['crypto']} dims {dims}", "raw_bytes": n * L, "ntt_bytes": ntt,
                   "import_from_hbm_ms": t_dev * 1e3, "import_from_host_ms": t_host * 1e3,
                   "save_ms": t_save * 1e3, "load_ms": t_load * 1e3,
                   "save_gbs": ntt / t_save / 1e9, "load_gbs": ntt / t_load / 1e9, "loaded_words_equal": same}
            log(f"dbfile {row['db']}: import {t_dev * 1e3:.1f} ms (HBM) / {t_host * 1e3:.1f} ms (host); "
                f"save {t_save * 1e3:.0f} ms ({row['save_gbs']:.2f} GB/s), load {t_load * 1e3:.0f} ms "
                f"({row['load_gbs']:.2f} GB/s), words equal {same}")
            rows.append(row)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    print(json.dumps({"workload": "dbfile", "rows": rows}), flush=True)
    ctx.close()


def load_gold():
    path = os.path.join(ROOT, "tests", "golden", "golden.json")
    return json.load(open(path)) if os.path.exists(path) else {}


def run_single(ctx, args, cfg, gold):
    """One database, one query: import, autotune, K replies back to back (value), the MAC
    kernel's HBM roofline, latency, the CPU baselines.  Returns the result dict."""
    import xpir_amd
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
    import_first_s = time.perf_counter() - t0
    # again, with the context's NTT tables for (n, m) already built: the per-database cost
    t0 = time.perf_counter()
    db2 = ctx.import_db_dev(prm, dbb, n, L, dims)
    ctx.sync()
    import_s = time.perf_counter() - t0
    db2.free()
    dbb.free()
    log(f"import (absorb + NTT) {import_s * 1e3:.1f} ms -> {raw / import_s / 1e9:.2f} GB/s raw "
        f"(first import, building the NTT tables: {import_first_s * 1e3:.1f} ms)")

    tuned = autotune_db(ctx, db, 0, args, cfg["name"])
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

    # throughput: K replies back to back on the context stream, one sync at each end; the
    # main MAC kernel is timed with HIP events inside this loop (events around it alone:
    # events around the whole level too would add launch gaps to every reply)
    ctx.timing(2)
    ctx.timing_read(reset=True)
    ctx.timing_read_main(reset=True)
    ctx.sync()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ctx.reply_dev(db, qd, rd)
    ctx.sync()
    total = time.perf_counter() - t_start
    main_ms, main_launches = ctx.timing_read_main(reset=True)
    # the whole MAC per reply (main kernel + scan), from a few more replies timed per level
    ctx.timing(True)
    ctx.timing_read(reset=True)
    mac_steps = max(1, min(args.steps, 5))
    for _ in range(mac_steps):
        ctx.reply_dev(db, qd, rd)
    ctx.sync()
    mac_ms, launches = ctx.timing_read(reset=True)
    ctx.timing_read_main(reset=True)
    ctx.timing(False)
    ms_per_step = total / args.steps * 1e3
    value = raw / (total / args.steps) / 1e9
    log(f"{args.steps} replies: {ms_per_step:.3f} ms/reply -> {value:.2f} GB/s")
    # latency: the same replies one at a time (submit -> reply in HBM, host-observed)
    lat = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        ctx.reply_dev(db, qd, rd)
        ctx.sync()
        lat.append(time.perf_counter() - t0)

    mn = prm.m * prm.n
    P = db.polys
    ntt_bytes = n * P * mn * 8
    q_bytes = qwords * 8
    reply_bytes = db.reply_words * 8
    alg_bytes = ntt_bytes + q_bytes + reply_bytes
    # the dominant kernel: the MAC's main kernel (by default the block-summary kernel), one
    # launch per reply, timed with HIP events on the context stream; it reads every DB and
    # query word once, and the reply words are written by the small scan kernel after it
    kern_name = main_kernel_name(tuned, dims[0])
    kern_s = main_ms / max(1, main_launches) / 1e3
    kern_bytes = ntt_bytes + q_bytes
    achieved = kern_bytes / kern_s / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(cfg["name"]),
                "traffic_from": "profiles/mac_traffic.json (calibrated FETCH_SIZE + WRITE_SIZE of this "
                                "kernel on this workload; not measured in this run)",
                "kernel": kern_name, "kernel_ms": main_ms / max(1, main_launches),
                "launches_per_reply": main_launches / args.steps, "alg_bytes": kern_bytes,
                "reply_mac_ms": mac_ms / mac_steps,
                "reply_mac_gbs": alg_bytes / (mac_ms / mac_steps / 1e3) / 1e9}

    host_io = None
    if args.host_io:
        # the boundary's host-buffer entry point (xpg_reply): query host -> HBM, reply,
        # reply HBM -> host; reported beside value, never as value
        q_host = qd.download()
        ctx.reply(db, q_host)  # staging buffer allocated outside the timed loop
        t0 = time.perf_counter()
        k = max(1, min(args.steps, 5))
        for _ in range(k):
            rep_h = ctx.reply(db, q_host)
        dt = (time.perf_counter() - t0) / k
        host_io = {"ms_per_reply": dt * 1e3, "gbs_raw_db": raw / dt / 1e9,
                   "query_gbs_over_pcie": q_host.nbytes / dt / 1e9,
                   "bit_exact": bool(np.array_equal(rep_h, reply)), "replies": k}
        log(f"host-buffer replies: {dt * 1e3:.1f} ms/reply -> {raw / dt / 1e9:.2f} GB/s raw (PCIe included)")
        del q_host

    cpu = cpu_all = None
    if not args.no_cpu_baseline:
        log("cpu baseline: the reference's MAC on this host over the same inputs")
        words = db.export_words()
        query = qd.download()
        cpu = cpu_baseline(words, query, cfg, prm, reply, raw)
        log(f"cpu baseline {cpu.get('value', 0):.3f} GB/s ({cpu.get('seconds', 0):.2f} s), "
            f"matches GPU reply: {cpu.get('matches_gpu_reply')}")
        cpu_all = cpu_all_cores(words, query, cfg, prm, reply, raw)
        log(f"cpu all cores {cpu_all['value']:.3f} GB/s on {cpu_all['cores']} threads "
            f"({cpu_all['seconds']:.2f} s), matches GPU reply: {cpu_all['matches_gpu_reply']}")

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
        "import_gbs": raw / import_s / 1e9, "import_first_ms": import_first_s * 1e3,
        "autotune": tuned,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_all_cores": cpu_all,
    }
    if host_io is not None:
        out["host_io"] = host_io
    for b in (skd, qd, rd):
        b.free()
    db.free()
    return out


def main_kernel_name(tuned, n_l):
    """The kernel a one-group level of n_l elements runs on: the tuning table's choice when
    the autotuner recorded one, else the fixed rule's (streaming block summaries, chunks
    of 8, for the single and plain workloads)."""
    for e in (tuned or {}).get("entries", []):
        if e["kind"] == 0 and e["n_l"] == n_l and e["groups"] == 1:
            k, a, b = e["kernel"], e["a"], e["b"]
            if k == 1:
                return "mac_direct_kernel"
            if k == 3:  # b: the load schedule (0 load then compute, 1 double-buffered, 2 rolling)
                return f"mac_stream_kernel<{a or 8},{b}>"
            return f"mac_batch2_kernel<{a},{b}>"
    return "mac_stream_kernel<8,0>"


def autotune_db(ctx, db, max_batch, args, label):
    """The MAC autotuner on db's level shapes (and batches up to max_batch queries) before
    any timed region; returns what it recorded and how long it took."""
    if args.no_autotune:
        return None
    if args.tune_file and not hasattr(args, "_tune_load"):
        args._tune_load = os.path.exists(args.tune_file)  # decided once, before any tuning
        if args._tune_load:
            ctx.tune_load(args.tune_file)
            log(f"tuning table loaded from {args.tune_file}")
    if getattr(args, "_tune_load", False):
        return {"seconds": 0.0, "entries": [], "loaded": args.tune_file}
    t0 = time.perf_counter()
    entries = ctx.autotune(db, max_batch=max_batch)
    dt = time.perf_counter() - t0
    if args.tune_file:
        ctx.tune_save(args.tune_file)
    names = {1: "one-chain", 3: "streaming", 4: "blocked"}
    for e in entries:
        what = f"level n_l={e['n_l']} groups={e['groups']} P={e['P']}" if e["kind"] == 0 else \
            f"batch of {e['qn']} on n_l={e['n_l']} groups={e['groups']} P={e['P']}"
        log(f"autotune {label} {what}: {names.get(e['kernel'], e['kernel'])}({e['a']},{e['b']},{e['c']}) "
            f"{e['ms']:.3f} ms vs {e['default_ms']:.3f} ms untuned")
    return {"seconds": dt, "entries": entries}


def _load_batch_dbs(ctx, scale, d_override=None, layout=0):
    """The eight uneven databases, imported on the GPU and resident (d = 2 grids), in the
    row-major (layout 0) or the matrix layout (1)."""
    import xpir_amd
    from xpir_amd.workloads import batch_databases
    dbs = []
    ctx.set_option("db_layout", layout)
    for spec in batch_databases(scale):
        dims = spec["dims"]
        prm = xpir_amd.params(spec["crypto"], dims[0])
        raw = spec["n_elems"] * spec["elem_bytes"]
        buf = ctx.alloc(raw + 16)
        ctx.random_bytes_dev(spec["db_seed"], 0, buf, raw)
        t0 = time.perf_counter()
        db = ctx.import_db_dev(prm, buf, spec["n_elems"], spec["elem_bytes"], dims)
        ctx.sync()
        dt = time.perf_counter() - t0
        buf.free()
        log(f"db {spec['n_elems']} x {spec['elem_bytes']} B dims {dims}: import {dt:.2f} s "
            f"({raw / dt / 1e9:.2f} GB/s raw), NTT {db.polys * dims[0] * dims[1] * prm.m * prm.n * 8 / 2**30:.2f} GiB")
        sk = xpir_amd.keygen(prm, spec["key_seed"])
        skd = ctx.alloc(sk.nbytes).upload(sk)
        dbs.append(dict(spec=spec, db=db, prm=prm, raw=raw, skd=skd, import_s=dt, matrix=layout == 1))
    ctx.set_option("db_layout", 0)
    return dbs


def bench_recursive(args):
    """The recursive configuration: LWE:100:4096:180, 2^20 x 1 KiB, d = 2 over 1024 x 1024
    (96 GiB NTT database resident).  Level 0 is 1024 groups of 1024 elements, each query
    word shared by 1024 groups, so it is bound by the integer issue rate rather than HBM;
    reported as exact Shoup steps per second.  The reply is checked by decrypting it on
    the GPU (xpg_decrypt_dev, level by level) back to the chosen element."""
    import xpir_amd
    from xpir_amd.workloads import C3_RECURSIVE as cfg
    ctx = xpir_amd.Context(0)
    prm = xpir_amd.params(cfg["crypto"], cfg["dims"][0])
    n, L, dims = cfg["n_elems"], cfg["elem_bytes"], cfg["dims"]
    raw = n * L
    dbb = ctx.alloc(raw + 16)
    ctx.random_bytes_dev(cfg["db_seed"], 0, dbb, raw)
    ctx.sync()
    t0 = time.perf_counter()
    db = ctx.import_db_dev(prm, dbb, n, L, dims)
    ctx.sync()
    import_s = time.perf_counter() - t0
    log(f"import of {n} x {L} B: {import_s:.3f} s")
    tuned = autotune_db(ctx, db, 0, args, "recursive")
    sk = xpir_amd.keygen(prm, cfg["key_seed"])
    skd = ctx.alloc(sk.nbytes).upload(sk)
    qd = ctx.alloc(db.query_words * 8)
    ctx.query_gen_dev(prm, skd, dims, cfg["index"], cfg["noise_seed"], qd)
    rd = ctx.alloc(db.reply_words * 8)
    for _ in range(max(1, min(args.warmup, 3))):
        ctx.reply_dev(db, qd, rd)
    ctx.sync()
    element = ctx.extract_reply_dev(prm, skd, rd.ptr, dims, L)
    want = np.empty(L, dtype=np.uint8)
    ctx._check(xpir_amd.capi.lib().xpg_d2h(ctx.h, want.ctypes.data, dbb.ptr + cfg["index"] * L, L))
    want = want.tobytes()
    dbb.free()
    ok = element == want
    log(f"reply decrypts to element {cfg['index']}: {ok}")
    steps = max(1, min(args.steps, 10))
    ctx.timing(True)
    ctx.timing_read(reset=True)
    ctx.timing_read_main(reset=True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.reply_dev(db, qd, rd)
    ctx.sync()
    dt = (time.perf_counter() - t0) / steps
    mac_ms, launches = ctx.timing_read(reset=True)
    ctx.timing(False)
    # exact steps per reply: chains x chain length, per level
    P0, P1 = db.polys, db.reply_polys  # level-0 polys per element; level-1 (final) polys
    mn2 = 2 * prm.m * prm.n
    steps_l0 = dims[1] * P0 * mn2 * dims[0]
    steps_l1 = P1 * mn2 * dims[1]
    out = {
        "metric": "DB GB/s per reply, batched vs plain, 1 MI355X, bit-exact; p50/p99 reply ms",
        "workload": "recursive", "value": raw / dt / 1e9, "unit": "GB/s", "ms_per_step": dt * 1e3,
        "steps": steps, "n_gpus": 1, "decrypts_to_element": bool(ok), "dtype": "u64",
        "data": "synthetic (seeded SplitMix64 bytes)",
        "config": {"workload": cfg["name"], "crypto": cfg["crypto"], "n_elems": n, "elem_bytes": L,
                   "dims": dims, "ntt_db_bytes": n * P0 * prm.m * prm.n * 8},
        "import_s": import_s, "mac_ms_per_reply": mac_ms / steps,
        "autotune": tuned,
        "shoup_steps_per_reply": steps_l0 + steps_l1,
        "shoup_steps_per_s": (steps_l0 + steps_l1) / (mac_ms / steps / 1e3),
    }
    print(json.dumps(out), flush=True)
    for b in (skd, qd, rd):
        b.free()
    db.free()
    ctx.close()


def bench_session(args):
    """Whole client sessions with the server program (xpir_amd/_build/pir_server) over a
    UNIX socket, driven by the reference client's messages (xpir_amd/pir_client.py): the
    single-query database (LWE:80:2048:120, 65536 x 4 KiB, one file that the server
    splits into elements, its -s option) queried at d = 1 (4 GiB query) and at d = 2 over
    256 x 256 (32 MiB query).  Per session: the client's upload time, the wait from its
    last query byte to the first reply byte, the download, and the server's own phase
    timings (upload, import overlapped with it, reply, send).  value = raw DB bytes /
    session time (first query byte sent .. last reply byte received) of the d = 1
    session.  Each reply is decrypted on the GPU back to the chosen element."""
    import re
    import shutil
    import socket
    import struct
    import subprocess
    import tempfile

    import xpir_amd
    from xpir_amd.pir_client import CLIENT_DRIVEN, PIRClient
    from xpir_amd.workloads import C2_SINGLE as cfg
    server = os.path.join(ROOT, "xpir_amd", "_build", "pir_server")
    n, L = cfg["n_elems"], cfg["elem_bytes"]
    raw = n * L
    ctx = xpir_amd.Context(0)
    dbb = ctx.alloc(raw + 16)
    ctx.random_bytes_dev(cfg["db_seed"], 0, dbb, raw)
    data = dbb.download(np.uint8, raw)
    dbb.free()
    tmp = tempfile.mkdtemp(prefix="xpir_session_")
    sessions = []
    try:
        os.makedirs(os.path.join(tmp, "db"))
        data.tofile(os.path.join(tmp, "db", "database.bin"))
        for dims in ([n], [256, 256]):
            prm = xpir_amd.params(cfg["crypto"], dims[0])
            index = cfg["index"] % n
            sk = xpir_amd.keygen(prm, cfg["key_seed"])
            skd = ctx.alloc(sk.nbytes).upload(sk)
            ct = 16 * prm.m * prm.n
            abs_b = prm.A * prm.n // 8
            qwords = sum(dims) * ct // 8
            qd = ctx.alloc(qwords * 8)
            ctx.query_gen_dev(prm, skd, dims, index, cfg["noise_seed"], qd)
            q = qd.download(np.uint64, qwords)
            qd.free()
            count = -(-L // abs_b)  # reply ciphertexts (computeReplySizeInChunks)
            for _ in range(1, len(dims)):
                count = -(-count * ct // abs_b)
            path = os.path.join(tmp, "pir.sock")
            lst = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
            lst.bind(path)
            lst.listen(1)
            logp = os.path.join(tmp, "server.log")
            with open(logp, "w") as logf:
                proc = subprocess.Popen([server, "--listen-fd", str(lst.fileno()), "--db-dir", os.path.join(tmp, "db"),
                                         "-s", str(n), "--sessions", "1"], pass_fds=[lst.fileno()], stdout=logf,
                                        stderr=subprocess.STDOUT)
            try:
                log(f"session dims {dims}: handshake")
                c = PIRClient(path, timeout=300)
                c.hello()
                kind, _, cnt, maxb = c.catalog()
                assert (cnt, maxb) == (n, L), (cnt, maxb)
                assert c.exchange_method() == CLIENT_DRIVEN
                c.optim_exit()
                c.send_crypto_params(cfg["crypto"] + ":?")
                c.send_key(struct.pack(f"<{prm.m}Q", *[prm.moduli[i] for i in range(prm.m)]))
                c.send_pir_params(len(dims), 1, dims)
                qb = memoryview(q.view(np.uint8))
                log(f"session dims {dims}: uploading {len(qb) / 2**20:.0f} MiB of query")
                t0 = time.perf_counter()
                for i in range(0, len(qb), ct):
                    c.send(qb[i:i + ct])
                t1 = time.perf_counter()
                first = c.recv(ct)
                t_first = time.perf_counter()
                rest = c.recv((count - 1) * ct)
                t2 = time.perf_counter()
                assert c.at_eof()
                c.close()
                rc = proc.wait(timeout=300)
            finally:
                if proc.poll() is None:
                    proc.kill()
                    proc.wait()
                lst.close()
                os.unlink(path)
            text = open(logp).read()
            if rc != 0 or "session done" not in text:
                raise RuntimeError(f"server failed (rc {rc}): {text}")
            m = re.search(r"query upload ([0-9.]+) ms, import ([0-9.]+) ms.*?, reply ([0-9.]+) ms, send ([0-9.]+) ms",
                          text)
            srv = {k: float(v) for k, v in zip(("upload_ms", "import_ms", "reply_ms", "send_ms"), m.groups())}
            reply = np.frombuffer(first + rest, dtype=np.uint64)
            rd = ctx.alloc(reply.nbytes).upload(reply)
            element = ctx.extract_reply_dev(prm, skd, rd.ptr, dims, L)
            rd.free()
            skd.free()
            ok = element == data[index * L:(index + 1) * L].tobytes()
            session_s = t2 - t0
            log(f"session dims {dims}: {session_s:.3f} s, server {srv}, decrypts to element: {ok}")
            sessions.append({"dims": dims, "query_bytes": len(qb), "reply_bytes": count * ct,
                             "client_upload_s": t1 - t0, "wait_first_reply_s": t_first - t1,
                             "download_s": t2 - t_first, "session_s": session_s,
                             "db_gbs": raw / session_s / 1e9, "server": srv, "decrypts_to_element": bool(ok)})
        concurrent = session_clients(ctx, server, tmp, cfg, data, n, L, clients=4)
        kept = session_kept_db(ctx, server, tmp, cfg, data, n, L)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
        ctx.close()
    out = {
        "metric": "session DB GB/s (first query byte sent .. last reply byte received), 1 MI355X, UNIX socket",
        "workload": "session", "value": sessions[0]["db_gbs"], "unit": "GB/s",
        "ms_per_step": sessions[0]["session_s"] * 1e3, "n_gpus": 1, "dtype": "u64",
        "data": "synthetic (seeded SplitMix64 bytes)",
        "config": {"workload": cfg["name"] + " via pir_server", "crypto": cfg["crypto"], "n_elems": n,
                   "elem_bytes": L},
        "sessions": sessions,
        "concurrent": concurrent,
        "kept_database": kept,
        "all_decrypt_to_element": (all(s["decrypts_to_element"] for s in sessions) and concurrent["all_decrypt"]
                                   and kept["decrypts_to_element"]),
    }
    print(json.dumps(out), flush=True)


def session_kept_db(ctx, server, tmp, cfg, data, n, L):
    """The reference's -z mode: a first client configures d = 2 over 256 x 256 and its
    imported database stays resident; a second client is then told the parameters and
    answered from it (no import in its session).  Times the second session."""
    import socket
    import struct
    import subprocess

    import xpir_amd
    from xpir_amd.pir_client import PIRClient
    dims = [256, 256]
    prm = xpir_amd.params(cfg["crypto"], dims[0])
    sk = xpir_amd.keygen(prm, cfg["key_seed"])
    skd = ctx.alloc(sk.nbytes).upload(sk)
    ct = 16 * prm.m * prm.n
    abs_b = prm.A * prm.n // 8
    count = -(-(-(-L // abs_b)) * ct // abs_b)  # two levels of reply ciphertexts
    key = struct.pack(f"<{prm.m}Q", *[prm.moduli[i] for i in range(prm.m)])
    qwords = sum(dims) * ct // 8

    def query(index, seed):
        qd = ctx.alloc(qwords * 8)
        ctx.query_gen_dev(prm, skd, dims, index, seed, qd)
        q = qd.download(np.uint64, qwords)
        qd.free()
        return memoryview(q.view(np.uint8))

    path = os.path.join(tmp, "pir_k.sock")
    lst = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
    lst.bind(path)
    lst.listen(2)
    logp = os.path.join(tmp, "server_k.log")
    with open(logp, "w") as logf:
        proc = subprocess.Popen([server, "--listen-fd", str(lst.fileno()), "--db-dir", os.path.join(tmp, "db"),
                                 "-s", str(n), "--sessions", "2", "-z", os.path.join(tmp, "PIRParams.cfg")],
                                pass_fds=[lst.fileno()], stdout=logf, stderr=subprocess.STDOUT)
    try:
        c = PIRClient(path, timeout=300)  # the configuring session
        c.hello()
        c.catalog()
        c.exchange_method()
        c.optim_exit()
        c.send_crypto_params(cfg["crypto"] + ":?")
        c.send_key(key)
        c.send_pir_params(2, 1, dims)
        qb = query(cfg["index"] % n, cfg["noise_seed"])
        for i in range(0, len(qb), ct):
            c.send(qb[i:i + ct])
        c.recv(count * ct)
        c.close()
        index = (cfg["index"] + 4321) % n
        qb = query(index, cfg["noise_seed"] + 1)
        log("kept database: second session")
        t0 = time.perf_counter()
        c = PIRClient(path, timeout=300)
        c.hello()
        c.catalog()
        c.exchange_method()
        c.recv_crypto_params()
        c.send_key(key)
        c.recv_pir_params()
        t1 = time.perf_counter()
        for i in range(0, len(qb), ct):
            c.send(qb[i:i + ct])
        reply = c.recv(count * ct)
        t2 = time.perf_counter()
        c.close()
        rc = proc.wait(timeout=120)
    finally:
        if proc.poll() is None:
            proc.kill()
            proc.wait()
        lst.close()
    rd = ctx.alloc(len(reply)).upload(np.frombuffer(reply, dtype=np.uint64))
    ok = ctx.extract_reply_dev(prm, skd, rd.ptr, dims, L) == data[index * L:(index + 1) * L].tobytes()
    rd.free()
    skd.free()
    text = open(logp).read().strip().splitlines()
    log(f"kept database: second session {t2 - t0:.3f} s (handshake {t1 - t0:.4f} s), server rc {rc}, "
        f"decrypts to element: {ok}; {text[-1] if text else ''}")
    return {"dims": dims, "session_s": t2 - t0, "handshake_s": t1 - t0, "query_and_reply_s": t2 - t1,
            "db_gbs": n * L / (t2 - t0) / 1e9, "server_rc": rc, "server_line": text[-1] if text else "",
            "decrypts_to_element": bool(ok)}


def session_clients(ctx, server, tmp, cfg, data, n, L, clients):
    """`clients` d = 1 sessions at once against pir_server --concurrent (each client on a
    thread of its own, each with its own query index): wall time from the first query
    byte sent to the last reply byte received, and the aggregate raw-DB rate."""
    import socket
    import struct
    import subprocess
    import threading

    import xpir_amd
    from xpir_amd.pir_client import PIRClient
    dims = [n]
    prm = xpir_amd.params(cfg["crypto"], n)
    sk = xpir_amd.keygen(prm, cfg["key_seed"])
    skd = ctx.alloc(sk.nbytes).upload(sk)
    ct = 16 * prm.m * prm.n
    qwords = n * ct // 8
    indices = [(cfg["index"] + 9973 * k) % n for k in range(clients)]
    queries = []
    for k, index in enumerate(indices):
        qd = ctx.alloc(qwords * 8)
        ctx.query_gen_dev(prm, skd, dims, index, cfg["noise_seed"] + k, qd)
        queries.append(qd.download(np.uint64, qwords))
        qd.free()
    path = os.path.join(tmp, "pir_c.sock")
    lst = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
    lst.bind(path)
    lst.listen(clients)
    logp = os.path.join(tmp, "server_c.log")
    with open(logp, "w") as logf:
        proc = subprocess.Popen([server, "--listen-fd", str(lst.fileno()), "--db-dir", os.path.join(tmp, "db"),
                                 "-s", str(n), "--sessions", str(clients), "--concurrent", str(clients)],
                                pass_fds=[lst.fileno()], stdout=logf, stderr=subprocess.STDOUT)
    replies, errors = [None] * clients, []
    try:
        conns = []
        for _ in range(clients):
            c = PIRClient(path, timeout=600)
            c.hello()
            c.catalog()
            c.exchange_method()
            c.optim_exit()
            c.send_crypto_params(cfg["crypto"] + ":?")
            c.send_key(struct.pack(f"<{prm.m}Q", *[prm.moduli[i] for i in range(prm.m)]))
            c.send_pir_params(1, 1, dims)
            conns.append(c)
        count = -(-L // (prm.A * prm.n // 8))

        def run(k):
            try:
                qb = memoryview(queries[k].view(np.uint8))
                for i in range(0, len(qb), ct):
                    conns[k].send(qb[i:i + ct])
                replies[k] = conns[k].recv(count * ct)
            except Exception as e:  # reported in the result
                errors.append(repr(e))

        log(f"{clients} concurrent d = 1 sessions: uploading {clients} x {qwords * 8 / 2**30:.0f} GiB")
        threads = [threading.Thread(target=run, args=(k,)) for k in range(clients)]
        t0 = time.perf_counter()
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        wall = time.perf_counter() - t0
        for c in conns:
            c.close()
        rc = proc.wait(timeout=300)
    finally:
        if proc.poll() is None:
            proc.kill()
            proc.wait()
        lst.close()
    ok = []
    for k, index in enumerate(indices):
        if replies[k] is None:
            ok.append(False)
            continue
        rep = np.frombuffer(replies[k], dtype=np.uint64)
        rd = ctx.alloc(rep.nbytes).upload(rep)
        ok.append(ctx.extract_reply_dev(prm, skd, rd.ptr, dims, L) == data[index * L:(index + 1) * L].tobytes())
        rd.free()
    skd.free()
    log(f"{clients} concurrent sessions: {wall:.3f} s wall, {clients * n * L / wall / 1e9:.2f} GB/s raw DB "
        f"aggregate, server rc {rc}, decrypt {ok}, errors {errors}")
    return {"clients": clients, "dims": dims, "wall_s": wall, "aggregate_db_gbs": clients * n * L / wall / 1e9,
            "query_gbs_aggregate": clients * qwords * 8 / wall / 1e9, "server_rc": rc, "errors": errors,
            "all_decrypt": bool(all(ok)) and not errors}


def bench_batched(args):
    """Batched uneven workload: every database's queries in one xpg_reply_batch_dev call,
    against the plain per-query path on the same databases."""
    import xpir_amd
    from xpir_amd.workloads import batch_query_counts
    ctx = xpir_amd.Context(0)
    dbs = _load_batch_dbs(ctx, args.scale, layout=args.db_layout)
    counts = batch_query_counts()
    rng = np.random.default_rng(42)
    reqs = []
    for i, (d, nq) in enumerate(zip(dbs, counts)):
        d["reqs"] = []
        for j in range(nq):
            qd = ctx.alloc(d["db"].query_words * 8)
            idx = int(rng.integers(0, d["spec"]["n_elems"]))
            ctx.query_gen_dev(d["prm"], d["skd"], d["spec"]["dims"], idx, 9000 + 100 * i + j, qd)
            rd = ctx.alloc(d["db"].reply_words * 8)
            d["reqs"].append((d["db"], qd, rd))
        reqs += d["reqs"]
    ctx.sync()
    log(f"{len(reqs)} queries over {len(dbs)} databases")
    tune = {}
    for i, (d, nq) in enumerate(zip(dbs, counts)):
        bucket = 1
        while bucket < min(nq, 64):
            bucket *= 2
        tune[i] = None if d.get("matrix") else autotune_db(ctx, d["db"], bucket if bucket >= 2 else 0, args,
                                                           f"db{i}")
    # plain: one query per database, alone (best of 2)
    plain_t = []
    for d in dbs:
        db, qd, rd = d["reqs"][0]
        best = 1e30
        for _ in range(2):
            ctx.sync()
            t0 = time.perf_counter()
            ctx.reply_dev(db, qd, rd)
            ctx.sync()
            best = min(best, time.perf_counter() - t0)
        d["plain_reply"] = rd.download()
        plain_t.append(best)
        log(f"plain reply {d['spec']['n_elems']} x {d['spec']['elem_bytes']} B: {best * 1e3:.1f} ms "
            f"({d['raw'] / best / 1e9:.2f} GB/s)")
    order = np.random.default_rng(43).permutation(len(reqs))
    batch = [reqs[k] for k in order]
    ctx.reply_batch_dev(batch)  # warm-up (allocates the batch buffers)
    ctx.sync()
    times = []
    for _ in range(max(1, args.steps)):
        ctx.sync()
        t0 = time.perf_counter()
        ctx.reply_batch_dev(batch)
        ctx.sync()
        times.append(time.perf_counter() - t0)
    makespan = min(times)
    exact = all(np.array_equal(d["reqs"][0][2].download(), d["plain_reply"]) for d in dbs)
    total_raw = sum(d["raw"] * nq for d, nq in zip(dbs, counts))
    plain_total = sum(t * nq for t, nq in zip(plain_t, counts))
    out = {
        "metric": "DB GB/s per reply, batched vs plain, 1 MI355X, bit-exact; p50/p99 reply ms",
        "workload": "batched", "value": total_raw / makespan / 1e9, "unit": "GB/s",
        "plain_value": total_raw / plain_total / 1e9, "batched_over_plain": plain_total / makespan,
        "makespan_ms": makespan * 1e3, "queries": int(sum(counts)), "n_gpus": 1,
        "bit_exact_vs_plain": bool(exact), "dtype": "u64", "data": "synthetic (seeded SplitMix64 bytes)",
        "config": {"workload": "8 uneven LWE:80:2048:120 databases, d=2, 1-64 queries each",
                   "scale": args.scale, "counts": counts,
                   "dbs": [[d["spec"]["n_elems"], d["spec"]["elem_bytes"], d["spec"]["dims"]] for d in dbs]},
        "plain_ms_per_db": [t * 1e3 for t in plain_t],
        "autotune_s": None if args.no_autotune else sum(v["seconds"] for v in tune.values()),
        "import_s": [d["import_s"] for d in dbs],
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_serving(args):
    """Serving: the batched databases resident, a seeded Poisson stream with Zipf(1.1)
    database choice through the continuous-batching server; replies checked against the
    same query answered alone."""
    import xpir_amd
    ctx = xpir_amd.Context(0)
    dbs = _load_batch_dbs(ctx, args.scale, layout=args.db_layout)
    # d = 1 views of the two smallest databases (their d = 1 queries fit the budget; the
    # absorption size depends on the first dimension, so each view is its own transform)
    for base in dbs[:2]:
        spec = dict(base["spec"])
        spec["dims"] = [spec["n_elems"]]
        prm = xpir_amd.params(spec["crypto"], spec["n_elems"])
        raw = spec["n_elems"] * spec["elem_bytes"]
        buf = ctx.alloc(raw + 16)
        ctx.random_bytes_dev(spec["db_seed"], 0, buf, raw)
        db = ctx.import_db_dev(prm, buf, spec["n_elems"], spec["elem_bytes"], spec["dims"])
        buf.free()
        sk = xpir_amd.keygen(prm, spec["key_seed"])
        base["d1"] = dict(spec=spec, db=db, prm=prm, raw=raw, skd=ctx.alloc(sk.nbytes).upload(sk))
    tune_s = 0.0
    for i, d in enumerate(dbs):
        for label, view in (("d2", d), ("d1", d.get("d1"))):
            if view is not None and not view.get("matrix"):  # the matrix layout has nothing to tune
                r = autotune_db(ctx, view["db"], 8, args, f"db{i}-{label}")
                tune_s += r["seconds"] if r else 0.0
    rng = np.random.default_rng(51)
    pool_n = 4
    for d in dbs[:2]:
        v = d["d1"]
        v["pool"] = []
        for j in range(2):
            qd = ctx.alloc(v["db"].query_words * 8)
            idx = int(rng.integers(0, v["spec"]["n_elems"]))
            ctx.query_gen_dev(v["prm"], v["skd"], v["spec"]["dims"], idx, 6000 + j, qd)
            rd = ctx.alloc(v["db"].reply_words * 8)
            ctx.reply_dev(v["db"], qd, rd)
            v["pool"].append((qd, hashlib.sha256(rd.download().tobytes()).hexdigest()))
            rd.free()
    for i, d in enumerate(dbs):
        d["pool"] = []
        for j in range(pool_n):
            qd = ctx.alloc(d["db"].query_words * 8)
            idx = int(rng.integers(0, d["spec"]["n_elems"]))
            ctx.query_gen_dev(d["prm"], d["skd"], d["spec"]["dims"], idx, 5000 + 100 * i + j, qd)
            rd = ctx.alloc(d["db"].reply_words * 8)
            ctx.reply_dev(d["db"], qd, rd)
            d["pool"].append((qd, hashlib.sha256(rd.download().tobytes()).hexdigest()))
            rd.free()
    ctx.sync()
    w = np.array([1.0 / (k + 1) ** 1.1 for k in range(len(dbs))])
    w /= w.sum()
    n = args.queries
    choice = rng.choice(len(dbs), size=n, p=w)
    pick = rng.integers(0, pool_n, size=n)
    use_d1 = rng.integers(0, 2, size=n)  # d in {1, 2} where a d = 1 view exists
    gaps = rng.exponential(1.0 / args.rate, size=n)
    srv = xpir_amd.Server(0, reply_pool_bytes=8 << 30, workers=args.workers, max_batch=64,
                          heavy_steps=args.heavy_steps)
    expected, lat, refused, served_raw, mismatches = {}, [], 0, 0, 0
    per_db = {}
    i = 0
    t0 = time.perf_counter()
    t_next = t0
    last_log = t0
    while i < n or expected:
        now = time.perf_counter()
        while i < n and now >= t_next:
            d = dbs[choice[i]]
            label = f"db{choice[i]}"
            if "d1" in d and use_d1[i]:
                d = d["d1"]
                label += "-d1"
            qd, digest = d["pool"][pick[i] % len(d["pool"])]
            t = srv.submit(d["db"], qd)
            if t is None:
                refused += 1
                break
            expected[t] = (digest, d["raw"], label)
            t_next += gaps[i]
            i += 1
        for c in srv.poll(timeout_us=500):
            digest, raw, label = expected.pop(c.ticket)
            got = ctx.download_ptr(c.reply_dev, c.reply_words)
            mismatches += hashlib.sha256(got.tobytes()).hexdigest() != digest
            lat.append((c.done_us - c.submit_us) / 1e3)
            per_db.setdefault(label, []).append(lat[-1])
            served_raw += raw
            srv.release(c.ticket)
        if now - last_log > 10:
            log(f"served {len(lat)}/{n}, in flight {len(expected)}, refused {refused}")
            last_log = now
    makespan = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.array(lat)
    out = {
        "metric": "DB GB/s per reply, batched vs plain, 1 MI355X, bit-exact; p50/p99 reply ms",
        "workload": "serving", "value": served_raw / makespan / 1e9, "unit": "GB/s",
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "latency_ms_per_db": {k: {"n": len(v), "p50": float(np.percentile(v, 50)), "p99": float(np.percentile(v, 99))}
                              for k, v in sorted(per_db.items())},
        "queries": n, "rate_per_s": args.rate, "makespan_s": makespan, "refused_submits": refused,
        "batches": st["batches"], "bit_exact": mismatches == 0, "mismatches": int(mismatches), "n_gpus": 1,
        "autotune_s": None if args.no_autotune else tune_s,
        "config": {"workload": "8 resident LWE:80:2048:120 databases, Zipf(1.1) choice, Poisson arrivals, "
                               "d=2 (d in {1,2} on the two smallest)",
                   "scale": args.scale, "workers": args.workers, "max_batch": 64, "reply_pool_bytes": 8 << 30,
                   "heavy_steps": args.heavy_steps, "db_layout": "matrix" if args.db_layout else "row-major"},
    }
    print(json.dumps(out), flush=True)
    ctx.close()


VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # lane ops/s: CUs x SIMDs x lanes x clock = 39.3 T


def batch_kernel_profile(fused: bool = True):
    """VALU lane instructions per exact chain step of the batched level-0 kernel (the fused
    persistent launch, or the per-chunk launches), and its HBM bytes per pass, from the
    rocprofv3 counter runs committed under profiles/ (the commands that made them are
    recorded there); None when absent."""
    name = "mac_fused_kernel" if fused else "mac_batch3_kernel"
    path = os.path.join(ROOT, "profiles", f"{name}_batched_uneven_8dbs.json")
    try:
        return json.load(open(path))
    except (OSError, ValueError):
        return None


MFMA_I8_PEAK_TOPS = 5000.0  # MI355X dense int8 matrix-core peak: 5.0 POP/s


def matrix_roofline(kernel_ms, launches, steps):
    """The matrix layout's level-0 kernel (mx_mac_kernel: the residues of the chains'
    prefixes) against the dense int8 matrix-core peak.  An exact step of a prefix is the 8 x
    8 byte-digit products of its two words: 64 multiply-adds, 128 int8 ops.  HBM traffic
    from the kernel's committed counter profile when present."""
    path = os.path.join(ROOT, "profiles", "mx_mac_kernel_batched_uneven_8dbs.json")
    try:
        prof = json.load(open(path))
    except (OSError, ValueError):
        prof = None
    achieved = steps * 128 / (kernel_ms / 1e3) / 1e12
    r = {"bound": "mfma", "unit": "TOP/s", "peak": MFMA_I8_PEAK_TOPS,
         "kernel": "mx_mac_kernel (int8 MFMA residues of the level-0 chains' prefixes, one launch per database "
                   "chunk)",
         "kernel_ms_per_pass": kernel_ms, "launches_per_pass": launches, "prefix_steps_per_pass": steps,
         "int8_ops_per_step": 128, "achieved": achieved, "frac": achieved / MFMA_I8_PEAK_TOPS,
         "peak_derivation": "AMD's published dense int8 MFMA rate of MI355X (5.0 POP/s, no sparsity)",
         "traffic": prof.get("hbm_bytes_per_pass") if prof else None}
    if prof:
        r.update(hbm_gbs=prof["hbm_bytes_per_pass"] / (kernel_ms / 1e3) / 1e9,
                 counters_from="profiles/mx_mac_kernel_batched_uneven_8dbs.json (rocprofv3 FETCH_SIZE, WRITE_SIZE "
                               "of the same pass; commands recorded there)")
    return r


def make_batch_queries(ctx, dbs, counts, rng):
    """Per database: its golden query first, then count-1 seeded queries (indices from
    rng, noise seeds 9000 + 100 i + j), all generated on the GPU."""
    from xpir_amd.workloads import batch_golden_query
    reqs = []
    for i, (d, nq) in enumerate(zip(dbs, counts)):
        d["reqs"] = []
        gq = batch_golden_query(i, d["spec"])
        for j in range(nq):
            idx = gq["index"] if j == 0 else int(rng.integers(0, d["spec"]["n_elems"]))
            seed = gq["noise_seed"] if j == 0 else 9000 + 100 * i + j
            qd = ctx.alloc(d["db"].query_words * 8)
            ctx.query_gen_dev(d["prm"], d["skd"], d["spec"]["dims"], idx, seed, qd)
            rd = ctx.alloc(d["db"].reply_words * 8)
            d["reqs"].append((d["db"], qd, rd))
        reqs += d["reqs"]
    ctx.sync()
    return reqs


def tune_batch_dbs(ctx, dbs, counts, args):
    tuned, secs = [], 0.0
    for i, (d, nq) in enumerate(zip(dbs, counts)):
        if d.get("matrix"):  # one level-0 kernel: nothing to tune
            tuned.append(None)
            continue
        bucket = 1
        while bucket < min(nq, 64):
            bucket *= 2
        r = autotune_db(ctx, d["db"], bucket if bucket >= 2 else 0, args, f"db{i}")
        tuned.append(r)
        secs += r["seconds"] if r else 0.0
    return tuned, secs


def bench_batched_pass(args):
    """One batched pass over the eight databases and nothing else on the GPU's MAC path
    after setup (the tuning table loaded from --tune-file, no tuning launches): the run
    that rocprofv3 profiles for the batched kernel's counters and its kernel trace."""
    import xpir_amd
    from xpir_amd.workloads import batch_query_counts
    ctx = xpir_amd.Context(0)
    dbs = _load_batch_dbs(ctx, args.scale, layout=args.db_layout)
    counts = batch_query_counts()
    reqs = make_batch_queries(ctx, dbs, counts, np.random.default_rng(42))
    if args.tune_file and os.path.exists(args.tune_file):
        ctx.tune_load(args.tune_file)
    order = np.random.default_rng(43).permutation(len(reqs))
    batch = [reqs[k] for k in order]
    ctx.set_option("batch_fuse", args.batch_fuse)
    ctx.timing(2)
    ctx.timing_read_batched(reset=True)
    passes = max(1, args.steps)
    for _ in range(passes):
        ctx.reply_batch_dev(batch)
    ctx.sync()
    ms, launches, steps = ctx.timing_read_batched(reset=True)
    print(json.dumps({"workload": "batched-pass", "batch_fuse": args.batch_fuse, "passes": passes,
                      "batch_kernel_ms_per_pass": ms / passes,
                      "batch_kernel_launches_per_pass": launches / passes,
                      "exact_steps_per_pass": steps / passes, "queries": len(reqs)}), flush=True)
    ctx.close()


def serving_latency(ctx, dbs, args):
    """p50/p99 reply latency (submit -> reply in HBM) of a seeded Poisson stream over the
    resident databases through the continuous-batching server: Zipf(1.1) database choice,
    d in {1, 2} (d = 1 through views of the two smallest databases), every reply compared
    with its query answered alone."""
    import xpir_amd
    from xpir_amd.workloads import serving_d1_views
    views = list(dbs)
    d1 = []
    for spec in serving_d1_views():
        prm = xpir_amd.params(spec["crypto"], spec["dims"][0])
        raw = spec["n_elems"] * spec["elem_bytes"]
        buf = ctx.alloc(raw + 16)
        ctx.random_bytes_dev(spec["db_seed"], 0, buf, raw)
        db = ctx.import_db_dev(prm, buf, spec["n_elems"], spec["elem_bytes"], spec["dims"])
        buf.free()
        sk = xpir_amd.keygen(prm, spec["key_seed"])
        d1.append(dict(spec=spec, prm=prm, db=db, raw=raw, skd=ctx.alloc(sk.nbytes).upload(sk)))
    views += d1
    rng = np.random.default_rng(51)
    bufs = []
    for vi, v in enumerate(views):
        v["pool"] = []
        for j in range(4):
            qd = ctx.alloc(v["db"].query_words * 8)
            ctx.query_gen_dev(v["prm"], v["skd"], v["spec"]["dims"], int(rng.integers(0, v["spec"]["n_elems"])),
                              5000 + 100 * vi + j, qd)
            rd = ctx.alloc(v["db"].reply_words * 8)
            ctx.reply_dev(v["db"], qd, rd)
            v["pool"].append((qd, rd.download()))
            bufs += [qd, rd]
    w = np.array([1.0 / (k + 1) ** 1.1 for k in range(8)])
    w /= w.sum()
    n = args.queries
    choice = rng.choice(8, size=n, p=w)
    use_d1 = rng.integers(0, 2, size=n)
    pick = rng.integers(0, 4, size=n)
    gaps = rng.exponential(1.0 / args.rate, size=n)
    srv = xpir_amd.Server(0, reply_pool_bytes=8 << 30, workers=2, max_batch=64)
    expected, lat, refused, served_raw, mismatches, i = {}, [], 0, 0, 0, 0
    t0 = t_next = last = time.perf_counter()
    while i < n or expected:
        now = time.perf_counter()
        while i < n and now >= t_next:
            v = views[8 + choice[i]] if choice[i] < 2 and use_d1[i] else views[choice[i]]
            qd, alone = v["pool"][pick[i]]
            t = srv.submit(v["db"], qd)
            if t is None:
                refused += 1
                break
            expected[t] = (alone, v["raw"])
            t_next += gaps[i]
            i += 1
        for c in srv.poll(timeout_us=200):
            alone, raw = expected.pop(c.ticket)
            mismatches += not np.array_equal(ctx.download_ptr(c.reply_dev, c.reply_words), alone)
            lat.append((c.done_us - c.submit_us) / 1e3)
            served_raw += raw
            srv.release(c.ticket)
        if now - last > 10:
            log(f"serving: {len(lat)}/{n} served, {len(expected)} in flight, {refused} refused")
            last = now
    makespan = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    for b in bufs:
        b.free()
    for v in d1:
        v["skd"].free()
        v["db"].free()
    lat = np.array(lat)
    return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
            "queries": n, "rate_per_s": args.rate, "served_gbs": served_raw / makespan / 1e9,
            "makespan_s": makespan, "batches": st["batches"], "refused_submits": refused,
            "bit_exact_vs_alone": mismatches == 0,
            "stream": "seeded Poisson arrivals, Zipf(1.1) over the 8 databases, d in {1,2} (d=1 views of the "
                      "two smallest), continuous-batching server: 2 workers, batches of <= 64, 8 GiB reply budget"}


def cpu_baseline_batched(dbs, counts, golden_replies, budget_steps=4.0e8):
    """The reference's reply arithmetic (oracle/_ref: its NFLlib deserializeDataNFL and
    mulandaddShoup, threads as its MULTI_THREAD build uses them: OpenMP over the polys of
    an element) on this host, for one query per database, on a bounded sample: level 0 on
    the first k groups of each grid (k groups cover ~budget_steps chain steps; the whole
    grid for the smallest database), scaled to all groups; the recursion import and the
    last level whole, as generateReply times them (PIRReplyGeneratorNFL_internal.cpp:359,444).
    value = raw bytes of the batch's replies / the summed CPU reply times."""
    from oracle.oracle import Oracle, RefNFL
    if not RefNFL.available():
        return {"error": "reference build (oracle/_ref) absent"}
    total_t, total_raw, cores, parts = 0.0, 0, 1, []
    for i, (d, nq) in enumerate(zip(dbs, counts)):
        spec, prm = d["spec"], d["prm"]
        d0, d1 = spec["dims"]
        orc = Oracle(spec["crypto"], d0)
        ref = RefNFL(prm.n, prm.m * 60)
        P0, mn = d["db"].polys, prm.m * prm.n
        q = d["golden_query"]
        q0, q1 = q[: d0 * 2 * mn], q[d0 * 2 * mn:]
        per_group = P0 * mn * 2 * d0
        k = d1 if i == 0 else max(1, min(d1, int(budget_steps // per_group)))
        words = ref.import_range(orc.A, spec["db_seed"], spec["n_elems"], spec["elem_bytes"], 0, k * d0, P0)
        rep0, t_l0 = ref.reply_level_timed(words, q0, d0, k, P0, P0)
        del words
        t_l0 *= d1 / k
        # the recursion import over all groups' level-0 ciphertexts (k sampled groups' real
        # words, the rest a stand-in of the same size: the import's cost is data-blind)
        ct = P0 * 2 * mn
        lvl0 = np.resize(rep0, d1 * ct)
        P1 = orc.next_polys(P0)
        nxt = np.zeros(d1 * P1 * mn, dtype=np.uint64)
        rb = lvl0.view(np.uint8)
        t1 = time.perf_counter()
        for g in range(d1):
            nxt[g * P1 * mn:(g + 1) * P1 * mn] = ref.import_buffers(rb[g * ct * 8:(g + 1) * ct * 8], P0, orc.A)
        t_rec = time.perf_counter() - t1
        final, t_l1 = ref.reply_level_timed(nxt, q1, d1, 1, P1, P1)
        t = t_l0 + t_rec + t_l1
        match = bool(np.array_equal(final, golden_replies[i])) if k == d1 else None
        parts.append({"db": i, "groups_sampled": k, "of_groups": d1, "reply_s": t, "level0_s": t_l0,
                      "recursion_s": t_rec, "level1_s": t_l1, "matches_gpu_reply": match,
                      "reply_sha256": hashlib.sha256(final.tobytes()).hexdigest() if k == d1 else None})
        total_t += t * nq
        total_raw += d["raw"] * nq
        cores = max(cores, P1)
        log(f"cpu baseline db{i}: {t:.3f} s per reply (level 0 from {k}/{d1} groups)")
    return {"value": total_raw / total_t / 1e9, "unit": "GB/s", "cores": cores, "kind": "reference",
            "seconds_per_batch": total_t, "per_db": parts,
            "sample": "one query per database: level 0 over the first k groups of its grid (all 100 for the "
                      "10000 x 256 B database, whose whole reply is compared with the GPU's), scaled to every "
                      "group; the recursion import and the last level whole; weighted by the batch's query "
                      "counts; threads = polys per element (the reference's OpenMP split)"}


def valu_roofline(args, k_ms, k_launches, k_steps, steps_s):
    """The row-major layout's level-0 kernel against the VALU issue rate, with the counters
    of its committed profile."""
    prof = batch_kernel_profile(bool(args.batch_fuse))
    roofline = {"bound": "valu", "unit": "TOP/s", "peak": VALU_PEAK_TOPS,
                "kernel": ("mac_fused_kernel (persistent register-blocked level-0 MAC over every database)"
                           if args.batch_fuse else
                           "mac_batch3_kernel (register-blocked level-0 MAC, every launch of the pass)"),
                "kernel_ms_per_pass": k_ms / args.steps, "launches_per_pass": k_launches / args.steps,
                "exact_steps_per_pass": k_steps / args.steps, "exact_steps_per_s": steps_s,
                "peak_derivation": "256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz = 39.3e12 32-bit lane ops/s "
                                   "(= 157.3 TFLOP/s fp32 / 2 flops per FMA / 2 for packed fp32)"}
    if prof:
        vps = prof["valu_lane_ops_per_step"]
        roofline.update(achieved=steps_s * vps / 1e12, frac=steps_s * vps / 1e12 / VALU_PEAK_TOPS,
                        valu_lane_ops_per_step=vps, traffic=prof.get("hbm_bytes_per_pass"),
                        hbm_gbs=(prof["hbm_bytes_per_pass"] / (k_ms / args.steps / 1e3) / 1e9
                                 if prof.get("hbm_bytes_per_pass") else None),
                        valu_busy=prof.get("valu_busy"), sustained_clock_ghz=prof.get("clock_ghz_from_grbm"),
                        counters_from=f"profiles/{'mac_fused_kernel' if args.batch_fuse else 'mac_batch3_kernel'}"
                                      "_batched_uneven_8dbs.json (rocprofv3 "
                                      "SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU, GRBM_GUI_ACTIVE, FETCH_SIZE, WRITE_SIZE "
                                      "of the same pass; commands recorded there; traffic is those counters', "
                                      "not measured in this run)")
    else:
        roofline.update(achieved=None, frac=None, traffic=None)
    return roofline


def bench_metric(args):
    """The metric: batched vs plain over the eight uneven databases, p50/p99 latency of a
    serving stream over them, bit-exactness against the reference's replies."""
    import xpir_amd
    from xpir_amd.workloads import C2_SINGLE, batch_query_counts
    gold = load_gold()
    ctx = xpir_amd.Context(0)
    t_setup = time.perf_counter()
    dbs = _load_batch_dbs(ctx, args.scale, layout=args.db_layout)
    counts = batch_query_counts()
    reqs = make_batch_queries(ctx, dbs, counts, np.random.default_rng(42))
    log(f"{len(reqs)} queries over {len(dbs)} databases, {sum(r[1].nbytes for r in reqs) / 2**30:.1f} GiB of queries "
        f"resident; setup {time.perf_counter() - t_setup:.1f} s")
    tuned, tune_s = tune_batch_dbs(ctx, dbs, counts, args)
    order = np.random.default_rng(43).permutation(len(reqs))
    batch = [reqs[k] for k in order]
    total_raw = sum(d["raw"] * nq for d, nq in zip(dbs, counts))

    # plain path: every query alone through xpg_reply_dev, one after another; per-query
    # latency (each synchronised) from a second pass
    plain = {}
    ctx.sync()
    t0 = time.perf_counter()
    for db, qd, rd in batch:
        ctx.reply_dev(db, qd, rd)
    ctx.sync()
    plain_s = time.perf_counter() - t0
    for db, qd, rd in reqs:
        plain[rd.ptr] = rd.download()
    plain_lat = []
    for db, qd, rd in batch:
        t0 = time.perf_counter()
        ctx.reply_dev(db, qd, rd)
        ctx.sync()
        plain_lat.append((time.perf_counter() - t0) * 1e3)
    log(f"plain path: {plain_s:.3f} s for {len(reqs)} queries -> {total_raw / plain_s / 1e9:.1f} GB/s")

    ctx.set_option("batch_fuse", args.batch_fuse)
    for _ in range(args.warmup):
        ctx.reply_batch_dev(batch)
    ctx.sync()
    ctx.timing(2)
    ctx.timing_read_batched(reset=True)
    ctx.timing_read_main(reset=True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.reply_batch_dev(batch)
    ctx.sync()
    total = time.perf_counter() - t0
    k_ms, k_launches, k_steps = ctx.timing_read_batched(reset=True)
    ctx.timing_read_main(reset=True)
    ctx.timing(False)
    pass_s = total / args.steps
    value = total_raw / pass_s / 1e9
    log(f"{args.steps} batched passes: {pass_s * 1e3:.1f} ms per pass -> {value:.1f} GB/s "
        f"({plain_s / pass_s:.2f}x the plain path), {k_launches / args.steps:.0f} level-0 launches per pass")
    # row-major layout: the other level-0 launch structure beside it, on a few passes (same
    # replies); the matrix layout has one launch structure (a launch per database chunk)
    other_launches = None
    if not args.db_layout:
        other = 1 - args.batch_fuse
        ctx.set_option("batch_fuse", other)
        ctx.reply_batch_dev(batch)
        ctx.timing(2)
        ctx.timing_read_batched(reset=True)
        ctx.sync()
        n_other = 3
        t0 = time.perf_counter()
        for _ in range(n_other):
            ctx.reply_batch_dev(batch)
        ctx.sync()
        other_s = (time.perf_counter() - t0) / n_other
        o_ms, o_launches, _ = ctx.timing_read_batched(reset=True)
        ctx.timing(False)
        ctx.set_option("batch_fuse", args.batch_fuse)
        other_launches = {"batch_fuse": other, "ms_per_pass": other_s * 1e3, "gbs": total_raw / other_s / 1e9,
                          "level0_kernel_ms_per_pass": o_ms / n_other,
                          "level0_launches_per_pass": o_launches / n_other, "passes": n_other}
        log(f"batch_fuse={other}: {other_s * 1e3:.1f} ms per pass, {o_launches / n_other:.0f} level-0 launches")

    # bit-exactness: every batched reply equals the plain reply; each database's golden
    # query equals the reference's reply bytes
    exact_vs_plain = all(np.array_equal(rd.download(), plain[rd.ptr]) for _, _, rd in reqs)
    golden_ok, golden_replies = [], []
    for i, d in enumerate(dbs):
        rep = d["reqs"][0][2].download()
        golden_replies.append(rep)
        g = gold.get("c4_batched", [None] * 8)[i]
        golden_ok.append(g is not None and hashlib.sha256(rep.tobytes()).hexdigest() == g["reply_sha256"])
        d["golden_query"] = d["reqs"][0][1].download()
    log(f"bit-exact: batched == plain {exact_vs_plain}, golden queries {sum(golden_ok)}/{len(golden_ok)}")

    steps_s = k_steps / (k_ms / 1e3)
    if args.db_layout:
        roofline = matrix_roofline(k_ms / args.steps, k_launches / args.steps, k_steps / args.steps)
    else:
        roofline = valu_roofline(args, k_ms, k_launches, k_steps, steps_s)
    latency = serving_latency(ctx, dbs, args)
    log(f"serving stream: p50 {latency['p50']:.2f} ms p99 {latency['p99']:.2f} ms, "
        f"{latency['served_gbs']:.1f} GB/s, exact {latency['bit_exact_vs_alone']}")

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline: the reference's reply arithmetic on this host, bounded sample")
        cpu = cpu_baseline_batched(dbs, counts, golden_replies)
    for _, qd, rd in reqs:
        qd.free()
        rd.free()
    for d in dbs:
        d["skd"].free()
        d["db"].free()

    single = None
    if not args.no_single:
        log("single-query configuration (65536 x 4 KiB) for the HBM-bound kernel")
        sargs = argparse.Namespace(**vars(args))
        sargs.steps, sargs.warmup = max(5, min(args.steps, 20)), 5
        sargs.host_io = True  # the host-buffer rate (PCIe included) rides along as a key of its own
        single = run_single(ctx, sargs, C2_SINGLE, gold.get("c2_single"))
        single = {k: single[k] for k in ("value", "ms_per_step", "bit_exact", "latency_ms", "import_gbs",
                                         "roofline", "cpu_baseline", "cpu_all_cores", "host_io", "config")
                  if k in single}
    ctx.close()
    plain_lat = np.array(plain_lat)
    out = {
        "metric": "DB GB/s per reply, batched vs plain, 1 MI355X, bit-exact; p50/p99 reply ms",
        "value": value, "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": pass_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic (seeded SplitMix64 bytes)",
        "config": {"workload": "batched uneven: 8 resident LWE:80:2048:120 databases, 10k..1M elements x "
                               "256 B..16 KiB, d=2 grids, 1-64 seeded queries each, one batched pass",
                   "crypto": "LWE:80:2048:120", "queries": len(reqs), "counts": counts,
                   "dbs": [[d["spec"]["n_elems"], d["spec"]["elem_bytes"], d["spec"]["dims"]] for d in dbs],
                   "raw_db_bytes_per_pass": total_raw},
        "plain_value": total_raw / plain_s / 1e9, "plain_ms_per_pass": plain_s * 1e3,
        "batched_over_plain": plain_s / pass_s,
        "bit_exact": bool(exact_vs_plain and all(golden_ok)),
        "bit_exact_detail": {"batched_vs_plain": bool(exact_vs_plain), "golden_queries": golden_ok},
        "latency_ms": {"p50": latency["p50"], "p99": latency["p99"], "source": "serving stream"},
        "serving": latency,
        "plain_latency_ms": {"p50": float(np.percentile(plain_lat, 50)), "p99": float(np.percentile(plain_lat, 99))},
        "autotune_s": tune_s,
        "db_layout": "matrix (level 0 on the int8 matrix cores)" if args.db_layout else "row-major",
        "level0_launch_structure": {"batch_fuse": None if args.db_layout else args.batch_fuse,
                                    "level0_launches_per_pass": k_launches / args.steps, "other": other_launches},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "single_query": single,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
