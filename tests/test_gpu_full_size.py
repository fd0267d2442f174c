"""The full-size configurations on the GPU against the reference's own reply bytes.

tests/golden/golden.json holds them (make_golden.py --full, from the reference's NFLlib
in oracle/_ref; the grids too large to hold as NTT databases on the host are
row-streamed one level-0 group at a time):

* recursive: LWE:100:4096:180, 2^20 x 1 KiB, d = 2 over 1024 x 1024 (A = 74, a 96 GiB
  NTT database, level-1 elements of 6 polys);
* batched uneven: the eight LWE:80:2048:120 databases (10 000 .. 1 000 000 elements of
  256 B .. 16 KiB, d = 2 grids from getDimSize), with the seeded query count of each,
  all in one xpg_reply_batch_dev call;
* serving: those eight databases resident plus d = 1 views of the two smallest, a seeded
  stream of 50 000 queries (Zipf(1.1) database choice, d in {1, 2}) through the
  continuous-batching server.

Bar: bit-exact.  Each golden query's reply equals the golden SHA-256; every other reply
equals the same query answered alone by the plain path; replies decrypt to the chosen
element.
"""
import hashlib
import json
import os
import time

import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _import(ctx, prm, spec):
    n, L = spec["n_elems"], spec["elem_bytes"]
    raw = ctx.alloc(n * L + 16)
    ctx.random_bytes_dev(spec["db_seed"], 0, raw, n * L)
    db = ctx.import_db_dev(prm, raw, n, L, spec["dims"])
    raw.free()
    return db


def _element(spec, index):
    orc = Oracle(spec["crypto"], spec["dims"][0])
    L = spec["elem_bytes"]
    return orc.rand_bytes(spec["db_seed"], index * L, L).tobytes()


def test_recursive_full_size_matches_golden(gpu_ctx):
    """2^20 x 1 KiB over 1024 x 1024: the reply equals the reference's (golden SHA-256,
    head and tail words) and decrypts to the chosen element."""
    import xpir_amd
    from xpir_amd.workloads import C3_RECURSIVE as cfg
    gold = GOLD["c3_recursive"]
    print("full size: recursive LWE:100:4096:180 2^20 x 1 KiB, 1024 x 1024", flush=True)
    prm = xpir_amd.params(cfg["crypto"], cfg["dims"][0])
    assert prm.A == gold["A"] == 74
    t0 = time.perf_counter()
    db = _import(gpu_ctx, prm, cfg)
    gpu_ctx.sync()
    print(f"import {time.perf_counter() - t0:.2f} s", flush=True)
    sk = xpir_amd.keygen(prm, cfg["key_seed"])
    skd = gpu_ctx.alloc(sk.nbytes).upload(sk)
    qd = gpu_ctx.alloc(db.query_words * 8)
    rd = gpu_ctx.alloc(db.reply_words * 8)
    try:
        gpu_ctx.query_gen_dev(prm, skd, cfg["dims"], cfg["index"], cfg["noise_seed"], qd)
        gpu_ctx.reply_dev(db, qd, rd)
        rep = rd.download()
        assert db.reply_polys == gold["Pd"] == 6
        assert [int(x) for x in rep[:16]] == gold["reply_head"]
        assert [int(x) for x in rep[-16:]] == gold["reply_tail"]
        assert _sha(rep) == gold["reply_sha256"]
        assert gpu_ctx.extract_reply_dev(prm, skd, rd.ptr, cfg["dims"], cfg["elem_bytes"]) == \
            _element(cfg, cfg["index"])
    finally:
        for b in (skd, qd, rd):
            b.free()
        db.free()


@pytest.fixture(scope="module", params=[0, 1], ids=["row_major", "matrix_layout"])
def batched(gpu_ctx, request):
    """The eight batched databases imported at full size and resident (109 GiB of NTT
    words), in the row-major layout or the matrix layout (option db_layout; one layout
    at a time on the GPU), with each one's secret key on the device."""
    import xpir_amd
    from xpir_amd.workloads import batch_databases
    out = []
    print(f"full size: importing the eight batched databases, db_layout {request.param}", flush=True)
    gpu_ctx.set_option("db_layout", request.param)
    try:
        for spec in batch_databases():
            prm = xpir_amd.params(spec["crypto"], spec["dims"][0])
            db = _import(gpu_ctx, prm, spec)
            sk = xpir_amd.keygen(prm, spec["key_seed"])
            out.append(dict(spec=spec, prm=prm, db=db, skd=gpu_ctx.alloc(sk.nbytes).upload(sk)))
    finally:
        gpu_ctx.set_option("db_layout", 0)
    gpu_ctx.sync()
    yield out
    for d in out:
        d["skd"].free()
        d["db"].free()


def _query(ctx, d, index, noise_seed):
    qd = ctx.alloc(d["db"].query_words * 8)
    ctx.query_gen_dev(d["prm"], d["skd"], d["spec"]["dims"], index, noise_seed, qd)
    return qd


def test_batched_uneven_full_size_matches_golden(gpu_ctx, batched):
    """Every database's seeded query count (334 queries) in one batched call, the
    databases interleaved: each database's golden query equals the reference's reply,
    every reply equals its query answered alone, golden replies decrypt."""
    from xpir_amd.workloads import batch_golden_query, batch_query_counts
    counts = batch_query_counts()
    rng = np.random.default_rng(42)
    reqs, golden_req = [], []
    try:
        for i, (d, nq) in enumerate(zip(batched, counts)):
            gq = batch_golden_query(i, d["spec"])
            for j in range(nq):
                if j == 0:
                    qd = _query(gpu_ctx, d, gq["index"], gq["noise_seed"])
                else:
                    qd = _query(gpu_ctx, d, int(rng.integers(0, d["spec"]["n_elems"])), 9000 + 100 * i + j)
                rd = gpu_ctx.alloc(d["db"].reply_words * 8)
                reqs.append((d["db"], qd, rd))
                if j == 0:
                    golden_req.append((i, d, qd, rd))
        print(f"full size: one batch of {len(reqs)} queries over 8 databases", flush=True)
        order = rng.permutation(len(reqs))
        t0 = time.perf_counter()
        gpu_ctx.reply_batch_dev([reqs[k] for k in order])
        gpu_ctx.sync()
        print(f"batched pass {time.perf_counter() - t0:.2f} s", flush=True)
        for i, d, qd, rd in golden_req:
            gold = GOLD["c4_batched"][i]
            assert d["prm"].A == gold["A"] and d["db"].reply_polys == gold["Pd"]
            assert _sha(rd.download()) == gold["reply_sha256"], f"database {i}"
            gq = batch_golden_query(i, d["spec"])
            assert gpu_ctx.extract_reply_dev(d["prm"], d["skd"], rd.ptr, d["spec"]["dims"],
                                             d["spec"]["elem_bytes"]) == _element(d["spec"], gq["index"])
        plain = {}
        for db, qd, rd in reqs:
            if id(db) not in plain:
                plain[id(db)] = gpu_ctx.alloc(db.reply_words * 8)
            gpu_ctx.reply_dev(db, qd, plain[id(db)])
            assert np.array_equal(rd.download(), plain[id(db)].download())
        for b in plain.values():
            b.free()
    finally:
        for _, qd, rd in reqs:
            qd.free()
            rd.free()


def test_serving_full_size_matches_golden(gpu_ctx, batched):
    """50 000 queries through the serving loop over the eight resident databases and
    d = 1 views of the two smallest: every served reply equals its query answered alone;
    the pools' golden queries answered alone equal the reference's replies."""
    import xpir_amd
    from xpir_amd.workloads import batch_golden_query, serving_d1_views
    views = []
    for i, d in enumerate(batched):
        gq = batch_golden_query(i, d["spec"])
        views.append(dict(d, gold=GOLD["c4_batched"][i], gq=gq))
    d1 = []
    for k, spec in enumerate(serving_d1_views()):
        prm = xpir_amd.params(spec["crypto"], spec["dims"][0])
        db = _import(gpu_ctx, prm, spec)
        sk = xpir_amd.keygen(prm, spec["key_seed"])
        d1.append(dict(spec=spec, prm=prm, db=db, skd=gpu_ctx.alloc(sk.nbytes).upload(sk),
                       gold=GOLD["c5_d1_views"][k], gq=dict(index=spec["index"], noise_seed=spec["noise_seed"])))
    views += d1
    rng = np.random.default_rng(51)
    bufs = []
    try:
        # a pool per view: the golden query and three more, each answered alone
        for vi, v in enumerate(views):
            v["pool"] = []
            for j in range(4):
                if j == 0:
                    index, seed = v["gq"]["index"], v["gq"]["noise_seed"]
                else:
                    index, seed = int(rng.integers(0, v["spec"]["n_elems"])), 5000 + 100 * vi + j
                qd = _query(gpu_ctx, v, index, seed)
                rd = gpu_ctx.alloc(v["db"].reply_words * 8)
                bufs += [qd, rd]
                gpu_ctx.reply_dev(v["db"], qd, rd)
                alone = rd.download()
                if j == 0:
                    assert _sha(alone) == v["gold"]["reply_sha256"], v["spec"]["n_elems"]
                    assert gpu_ctx.extract_reply_dev(v["prm"], v["skd"], rd.ptr, v["spec"]["dims"],
                                                     v["spec"]["elem_bytes"]) == _element(v["spec"], index)
                v["pool"].append((qd, alone))
        # Zipf(1.1) over the eight databases; d = 1 or 2 (coin) where a d = 1 view exists
        w = np.array([1.0 / (k + 1) ** 1.1 for k in range(8)])
        w /= w.sum()
        n = 50_000
        choice = rng.choice(8, size=n, p=w)
        use_d1 = rng.integers(0, 2, size=n)
        pick = rng.integers(0, 4, size=n)
        srv = xpir_amd.Server(0, reply_pool_bytes=8 << 30, workers=2, max_batch=64)
        print(f"full size: serving {n} queries", flush=True)
        try:
            expected, mismatches, served, refused, i = {}, 0, 0, 0, 0
            t0 = last = time.perf_counter()
            while i < n or expected:
                while i < n:
                    v = views[8 + choice[i]] if choice[i] < 2 and use_d1[i] else views[choice[i]]
                    qd, alone = v["pool"][pick[i]]
                    t = srv.submit(v["db"], qd)
                    if t is None:
                        refused += 1
                        break
                    expected[t] = alone
                    i += 1
                for c in srv.poll(timeout_us=500):
                    assert c.status == 0
                    alone = expected.pop(c.ticket)
                    mismatches += not np.array_equal(gpu_ctx.download_ptr(c.reply_dev, c.reply_words), alone)
                    served += 1
                    srv.release(c.ticket)
                now = time.perf_counter()
                if now - last > 15:
                    print(f"served {served}/{n}", flush=True)
                    last = now
                assert now - t0 < 600, "serving stream did not drain"
            st = srv.stats()
        finally:
            srv.close()
        print(f"served {served} in {time.perf_counter() - t0:.1f} s, {st['batches']} batches, "
              f"{refused} submits refused", flush=True)
        assert served == n and mismatches == 0
        assert st["batches"] < n
    finally:
        for b in bufs:
            b.free()
        for v in d1:
            v["skd"].free()
            v["db"].free()
