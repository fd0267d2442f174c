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
    # budget: room for a handful of replies at once, so admission refuses some
    srv = xpir_amd.Server(0, reply_pool_bytes=6 << 20, workers=2, max_batch=16)
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
