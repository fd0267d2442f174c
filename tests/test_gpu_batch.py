"""Batched replies (xpg_reply_batch_dev): many queries over several resident databases
of uneven element count and size, interleaved.  Every reply must equal the plain
single-query path's reply and the oracle's (bit-exact), whatever the mix."""
import numpy as np
import pytest

from oracle.oracle import Oracle, extract_reply

pytestmark = pytest.mark.gpu

C2 = "LWE:80:2048:120"


def _setup_db(ctx, orc, prm, n, L, dims, seed):
    data = orc.rand_bytes(seed, 0, n * L)
    db = ctx.import_db(prm, data, n, L, dims)
    return db, data


@pytest.mark.parametrize("specs", [
    # (n_elems, elem_bytes, dims, n_queries)
    [(60, 256, [10, 6], 5), (35, 1024, [7, 5], 3), (48, 4096, [8, 6], 9)],
    [(400, 300, [20, 20], 12)],
    [(90, 512, [90], 4), (30, 2000, [30], 2)],       # d = 1 batches
    [(96, 200, [8, 4, 3], 20)],                       # d = 3: levels 1 and 2 grouped over requests
])
@pytest.mark.parametrize("budget", [0, 1])             # 1 byte: one request per level group
def test_batch_equals_plain_and_oracle(gpu_ctx, specs, budget):
    import xpir_amd
    gpu_ctx.set_option("group_budget", budget)
    rng = np.random.default_rng(len(specs))
    dbs, reqs, expect = [], [], []
    for si, (n, L, dims, nq) in enumerate(specs):
        orc = Oracle(C2, dims[0])
        prm = xpir_amd.params(C2, dims[0])
        db, data = _setup_db(gpu_ctx, orc, prm, n, L, dims, 100 + si)
        dbs.append(db)
        sk = orc.keygen(40 + si)
        for qi in range(nq):
            index = int(rng.integers(0, n))
            q = np.concatenate(orc.make_query(sk, dims, index, 1000 * si + qi))
            qd = gpu_ctx.alloc(q.nbytes).upload(q)
            rd = gpu_ctx.alloc(db.reply_words * 8)
            reqs.append((db, qd, rd))
            expect.append((orc, sk, dims, L, data[index * L:(index + 1) * L].tobytes(), q, db))
    # interleave databases in the request list
    perm = rng.permutation(len(reqs))
    batch = [reqs[i] for i in perm]
    gpu_ctx.reply_batch_dev(batch)
    for i in perm:
        db, qd, rd = reqs[i]
        orc, sk, dims, L, element, q, _ = expect[i]
        got = rd.download()
        plain = gpu_ctx.reply(db, q)
        assert np.array_equal(got, plain)
        assert extract_reply(orc, sk, got, dims, L) == element
    # the oracle on the first request of each database
    seen = set()
    for i in range(len(reqs)):
        db, qd, rd = reqs[i]
        if id(db) in seen:
            continue
        seen.add(id(db))
        orc, sk, dims, L, element, q, _ = expect[i]
        words = db.export_words()
        levels, off = [], 0
        for nl in dims:
            levels.append(q[off:off + nl * 2 * orc.mn])
            off += nl * 2 * orc.mn
        want, _ = orc.generate_reply(words, db.polys, dims, levels)
        assert np.array_equal(rd.download(), want)
    for db, qd, rd in reqs:
        qd.free()
        rd.free()
    for db in dbs:
        db.free()
    gpu_ctx.set_option("group_budget", 0)
