"""Serving loop (xpg_server_*): a seeded Poisson stream over several resident databases
with Zipf database choice, d in {1, 2}, admission by reply memory (a budget small enough
to refuse some submissions), continuous batching on two workers.  Every reply must be
the reply of the same query computed alone (plain path), bit for bit; every reply
computed alone equals the oracle's and decrypts to the chosen element."""
import time

import numpy as np
import pytest

from oracle.oracle import Oracle, extract_reply

pytestmark = pytest.mark.gpu

C2 = "LWE:80:2048:120"


def test_serving_mix_equals_alone(gpu_ctx):
    import xpir_amd
    rng = np.random.default_rng(7)
    specs = [(200, 256, [20, 10]), (64, 1024, [64]), (150, 4096, [15, 10]), (40, 2048, [40])]
    dbs = []
    for si, (n, L, dims) in enumerate(specs):
        orc = Oracle(C2, dims[0])
        prm = xpir_amd.params(C2, dims[0])
        data = orc.rand_bytes(500 + si, 0, n * L)
        db = gpu_ctx.import_db(prm, data, n, L, dims)
        P = orc.polys(L)
        words = np.zeros(int(np.prod(dims)) * P * orc.mn, dtype=np.uint64)
        words[: n * P * orc.mn] = np.concatenate([orc.import_element(data[e * L:(e + 1) * L]) for e in range(n)])
        sk = orc.keygen(600 + si)
        # a pool of distinct queries per database, each with its reply computed alone
        pool = []
        for qi in range(6):
            index = int(rng.integers(0, n))
            qlevels = orc.make_query(sk, dims, index, 7000 + 10 * si + qi)
            q = np.concatenate(qlevels)
            qd = gpu_ctx.alloc(q.nbytes).upload(q)
            alone = gpu_ctx.reply(db, q)
            assert np.array_equal(alone, orc.generate_reply(words, P, dims, qlevels)[0])
            assert extract_reply(orc, sk, alone, dims, L) == data[index * L:(index + 1) * L].tobytes()
            pool.append((qd, alone))
        dbs.append((db, pool))
    # Zipf(1.1) database choice, Poisson arrivals
    w = np.array([1.0 / (i + 1) ** 1.1 for i in range(len(dbs))])
    w /= w.sum()
    n_req = 300
    choice = rng.choice(len(dbs), size=n_req, p=w)
    qpick = rng.integers(0, 6, size=n_req)
    # bursts: Poisson arrivals at 4000/s, with every 25th gap stretched so requests
    # pile up behind a running batch and must share the next one
    gaps = rng.exponential(1.0 / 4000.0, size=n_req)
    gaps[np.arange(n_req) % 25 != 24] = 0.0
    # pool: 8 MiB of reply slots (about twenty replies at once, so bursts are refused) and
    # two 4 MiB workspaces (a d = 2 request here needs 2.6 MiB, so batches stay small)
    srv = xpir_amd.Server(0, reply_pool_bytes=16 << 20, workers=2, max_batch=16)
    try:
        expected, done, refused = {}, {}, 0
        t0 = time.perf_counter()
        t_next = t0
        i = 0
        while i < n_req or len(done) < n_req:
            now = time.perf_counter()
            while i < n_req and now >= t_next:
                db, pool = dbs[choice[i]]
                qd, alone = pool[qpick[i]]
                t = srv.submit(db, qd)
                if t is None:
                    refused += 1
                    break
                expected[t] = alone
                t_next += gaps[i]
                i += 1
            for c in srv.poll(timeout_us=200 if i >= n_req or time.perf_counter() < t_next else 0):
                assert c.status == 0
                got = gpu_ctx.download_ptr(c.reply_dev, c.reply_words)
                assert np.array_equal(got, expected[c.ticket]), f"ticket {c.ticket}"
                done[c.ticket] = (c.submit_us, c.done_us)
                srv.release(c.ticket)
            assert time.perf_counter() - t0 < 120, "serving stream did not drain"
        st = srv.stats()
        lat = np.array([d - s for s, d in done.values()]) / 1e3
        print(f"served {len(done)} replies, {st['batches']} batches, refused {refused}, "
              f"p50 {np.percentile(lat, 50):.2f} ms p99 {np.percentile(lat, 99):.2f} ms")
        assert st["batched_requests"] == n_req
        assert st["batches"] < n_req  # some requests shared a batch
    finally:
        srv.close()
        for db, pool in dbs:
            for qd, _ in pool:
                qd.free()
            db.free()


def _hip_free_bytes():
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_admission_bounds_device_memory(gpu_ctx):
    """The server's device memory is its pool: a request whose intermediates exceed a
    worker's workspace is refused (XPG_ESIZE), bursts of requests that fit are served in
    batches sized to the workspace with replies equal to the plain path's, device memory
    stays within the pool (plus the workers' NTT tables and streams), and a ticket still
    in flight keeps its slot (release gives XPG_EBUSY)."""
    import xpir_amd
    from xpir_amd.capi import XpgError
    C2 = "LWE:80:2048:120"
    # level-0 intermediates per request: 2048 groups x 64 KiB = 128 MiB, above a 16 MiB workspace
    big_prm = xpir_amd.params(C2, 8)
    raw = gpu_ctx.alloc(8 * 2048 * 64 + 16)
    gpu_ctx.random_bytes_dev(71, 0, raw, 8 * 2048 * 64)
    big = gpu_ctx.import_db_dev(big_prm, raw, 8 * 2048, 64, [8, 2048])
    raw.free()
    # per request: 16 groups x 64 KiB of level-0 outputs, plus 3 MiB of level-1 database
    # for the batch: up to 13 requests per 16 MiB workspace
    orc = Oracle(C2, 64)
    prm = xpir_amd.params(C2, 64)
    n, L, dims = 1000, 300, [64, 16]
    data = orc.rand_bytes(72, 0, n * L)
    small = gpu_ctx.import_db(prm, data, n, L, dims)
    sk = orc.keygen(73)
    pool = []
    for qi in range(4):
        q = np.concatenate(orc.make_query(sk, dims, 37 * qi + 5, 740 + qi))
        pool.append((gpu_ctx.alloc(q.nbytes).upload(q), gpu_ctx.reply(small, q)))
    bq = gpu_ctx.alloc(big.query_words * 8)
    gpu_ctx.sync()
    free0 = _hip_free_bytes()
    budget = 64 << 20  # 32 MiB of reply slots (85 replies of 384 KiB), two 16 MiB workspaces
    srv = xpir_amd.Server(0, reply_pool_bytes=budget, workers=2, max_batch=64)
    try:
        with pytest.raises(XpgError, match="-4"):
            srv.submit(big, bq)
        lowest, busy_seen, done = free0, False, 0
        for burst in range(3):
            expected = {}
            for k in range(40):  # back to back: they pile up behind the running batches
                qd, alone = pool[k % 4]
                t = srv.submit(small, qd)
                assert t is not None
                expected[t] = alone
            if burst == 0:
                try:
                    srv.release(t)  # the burst's last request: still pending or running
                except XpgError as e:
                    busy_seen = "-5" in str(e)
            lowest = min(lowest, _hip_free_bytes())
            t0 = time.perf_counter()
            while expected:
                for c in srv.poll(timeout_us=2000):
                    assert c.status == 0
                    assert np.array_equal(gpu_ctx.download_ptr(c.reply_dev, c.reply_words), expected.pop(c.ticket))
                    srv.release(c.ticket)
                    done += 1
                lowest = min(lowest, _hip_free_bytes())
                assert time.perf_counter() - t0 < 60
        st = srv.stats()
        print(f"{st['batches']} batches for {done} requests; device memory used {(free0 - lowest) >> 20} MiB "
              f"for a {budget >> 20} MiB pool")
        assert busy_seen and done == 120
        assert st["batches"] < 120
        assert free0 - lowest <= budget + (32 << 20)  # the pool, the workers' tables and streams
    finally:
        srv.close()
        for qd, _ in pool:
            qd.free()
        bq.free()
        big.free()
        small.free()


def test_concurrent_submitters(gpu_ctx):
    """Four host threads submit at once (the serving API takes submits from any thread)
    while the main thread polls and releases: every reply equals its query answered
    alone, every ticket is unique, and nothing is lost."""
    import threading

    import xpir_amd
    specs = [(300, 512, [30, 10]), (64, 2048, [64])]
    dbs = []
    for si, (n, L, dims) in enumerate(specs):
        orc = Oracle(C2, dims[0])
        prm = xpir_amd.params(C2, dims[0])
        db = gpu_ctx.import_db(prm, orc.rand_bytes(900 + si, 0, n * L), n, L, dims)
        sk = orc.keygen(910 + si)
        pool = []
        for qi in range(3):
            q = np.concatenate(orc.make_query(sk, dims, 17 * qi + si, 920 + 10 * si + qi))
            pool.append((gpu_ctx.alloc(q.nbytes).upload(q), gpu_ctx.reply(db, q)))
        dbs.append((db, pool))
    srv = xpir_amd.Server(0, reply_pool_bytes=64 << 20, workers=2, max_batch=16)
    expected, lock, errors = {}, threading.Lock(), []
    per_thread = 60

    def submitter(tid):
        try:
            rng = np.random.default_rng(tid)
            k = 0
            while k < per_thread:
                db, pool = dbs[int(rng.integers(0, 2))]
                qd, alone = pool[int(rng.integers(0, 3))]
                t = srv.submit(db, qd)
                if t is None:  # reply slots full: the main thread releases, try again
                    time.sleep(0.0005)
                    continue
                with lock:
                    assert t not in expected
                    expected[t] = alone
                k += 1
        except Exception as e:  # surfaced by the main thread
            errors.append(e)

    try:
        threads = [threading.Thread(target=submitter, args=(i,)) for i in range(4)]
        for th in threads:
            th.start()
        done, t0, early = 0, time.perf_counter(), []
        while done < 4 * per_thread:
            # a completion can arrive before its submitter has recorded the ticket: keep it
            # until the ticket appears
            for c in early + srv.poll(timeout_us=1000):
                assert c.status == 0
                with lock:
                    alone = expected.pop(c.ticket, None)
                if alone is None:
                    early.append(c)
                    continue
                if c in early:
                    early.remove(c)
                assert np.array_equal(gpu_ctx.download_ptr(c.reply_dev, c.reply_words), alone)
                srv.release(c.ticket)
                done += 1
            assert not errors, errors
            assert time.perf_counter() - t0 < 60, f"{done} of {4 * per_thread} served"
        for th in threads:
            th.join()
        assert not expected and srv.stats()["batched_requests"] == 4 * per_thread
    finally:
        srv.close()
        for db, pool in dbs:
            for qd, _ in pool:
                qd.free()
            db.free()


def test_heavy_batches_run_one_at_a_time(gpu_ctx):
    """heavy_steps: a burst over a heavy and a light database on two workers.  Batches of
    the heavy database (each above the threshold) never run at the same time -- each
    starts after the previous one's completion -- while light batches still run; every
    reply equals the same query answered alone."""
    import xpir_amd
    rng = np.random.default_rng(17)
    # (n, L, dims): about 18.7 M exact steps per heavy request, 2.2 M per light one
    specs = [(2000, 4096, [50, 40]), (200, 256, [20, 10])]
    dbs = []
    for si, (n, L, dims) in enumerate(specs):
        orc = Oracle(C2, dims[0])
        prm = xpir_amd.params(C2, dims[0])
        db = gpu_ctx.import_db(prm, orc.rand_bytes(800 + si, 0, n * L), n, L, dims)
        sk = orc.keygen(810 + si)
        pool = []
        for qi in range(4):
            q = np.concatenate(orc.make_query(sk, dims, int(rng.integers(0, n)), 8200 + 10 * si + qi))
            pool.append((gpu_ctx.alloc(q.nbytes).upload(q), gpu_ctx.reply(db, q)))
        dbs.append((db, pool))
    srv = xpir_amd.Server(0, reply_pool_bytes=256 << 20, workers=2, max_batch=4, heavy_steps=10_000_000)
    try:
        expected, batches = {}, {}
        for i in range(60):  # heavy and light requests interleaved, all at once
            db, pool = dbs[i % 2]
            qd, alone = pool[(i // 2) % 4]
            t = srv.submit(db, qd)
            assert t is not None
            expected[t] = (i % 2, alone)
        t0 = time.perf_counter()
        while expected:
            for c in srv.poll(timeout_us=1000):
                assert c.status == 0
                kind, alone = expected.pop(c.ticket)
                assert np.array_equal(gpu_ctx.download_ptr(c.reply_dev, c.reply_words), alone)
                batches.setdefault((kind, c.start_us), c.done_us)
                srv.release(c.ticket)
            assert time.perf_counter() - t0 < 60, "burst did not drain"
        heavy = sorted((s, d) for (k, s), d in batches.items() if k == 0)
        light = [s for (k, s) in batches if k == 1]
        print(f"{len(heavy)} heavy batches, {len(light)} light batches", flush=True)
        assert len(heavy) >= 8 and light
        for (s0, d0), (s1, _) in zip(heavy, heavy[1:]):
            assert s1 >= d0, "two heavy batches ran at the same time"
    finally:
        srv.close()
        for db, pool in dbs:
            for qd, _ in pool:
                qd.free()
            db.free()
