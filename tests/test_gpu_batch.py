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
@pytest.mark.parametrize("fuse", [1, 0])               # level 0 fused over the databases, or per chunk
def test_batch_equals_plain_and_oracle(gpu_ctx, specs, budget, fuse):
    import xpir_amd
    gpu_ctx.set_option("group_budget", budget)
    gpu_ctx.set_option("batch_fuse", fuse)
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
    gpu_ctx.set_option("batch_fuse", 1)


SHAPES = [(2, g, q) for g, q in ((1, 1), (1, 2), (1, 4), (4, 1), (4, 2), (4, 4), (8, 1), (8, 2), (16, 1))] + \
         [(3, g, q) for g, q in ((4, 1), (4, 2), (4, 4), (8, 1), (8, 2))]


# 7, 6, 5: a ragged last query block of 3, 2, 1 queries takes a launch of its own
@pytest.mark.parametrize("qn", [8, 7, 6, 5])
def test_every_register_block_shape_equals_plain(gpu_ctx, qn):
    """Each built (family, rows, queries) shape of the level-0 batch kernels, forced
    through a tuning entry, gives every query's plain reply (family 3 stages the query
    words through LDS for a workgroup's four waves).  Per-chunk launches (batch_fuse 0):
    the tuning entry picks their shape."""
    import xpir_amd
    gpu_ctx.set_option("batch_fuse", 0)
    n, L, dims = 600, 256, [30, 20]
    orc = Oracle(C2, dims[0])
    prm = xpir_amd.params(C2, dims[0])
    db, _ = _setup_db(gpu_ctx, orc, prm, n, L, dims, 77)
    sk = orc.keygen(78)
    rng = np.random.default_rng(79)
    reqs, plain = [], []
    try:
        for qi in range(qn):
            q = np.concatenate(orc.make_query(sk, dims, int(rng.integers(0, n)), 800 + qi))
            reqs.append((db, gpu_ctx.alloc(q.nbytes).upload(q), gpu_ctx.alloc(db.reply_words * 8)))
            plain.append(gpu_ctx.reply(db, q))
        for fam, gr, qr in SHAPES:
            gpu_ctx.tune_clear()
            gpu_ctx.tune_set([{"kind": 1, "n": prm.n, "m": prm.m, "P": db.polys, "n_l": dims[0], "groups": dims[1],
                               "qn": 8, "kernel": 4, "a": gr, "b": qr, "c": fam, "ms": 1.0, "default_ms": 1.0}])
            for _, _, rd in reqs:
                rd.upload(np.zeros(db.reply_words, dtype=np.uint64))
            gpu_ctx.reply_batch_dev(reqs)
            for k, (_, _, rd) in enumerate(reqs):
                assert np.array_equal(rd.download(), plain[k]), (fam, gr, qr, k)
    finally:
        gpu_ctx.tune_clear()
        gpu_ctx.set_option("batch_fuse", 1)
        for _, qd, rd in reqs:
            qd.free()
            rd.free()
        db.free()


# (n_elems, elem_bytes, dims, queries): every level 0 but one has enough rows (groups x
# polys) for its chunk to run register-blocked and join a fused launch; the one with a
# single query is answered alone.  Query counts give blocks of four plus ragged blocks of
# 1, 2 and 3; 13 000 B elements take two polys; the d = 1 database has one group of 4
# polys (its only row block is mostly rows past the level); one database has d = 3.
# Fourteen fused chunks make 26 segments: two launches.
FUSED_DBS = [
    (400, 256, [20, 20], 5), (300, 13000, [15, 20], 7), (512, 1024, [16, 32], 5), (180, 600, [10, 18], 9),
    (96, 200, [8, 4, 3], 9), (40, 40000, [40], 41), (250, 300, [25, 20], 2), (160, 4096, [10, 16], 3),
    (330, 128, [30, 11], 10), (200, 5000, [10, 20], 6), (64, 2048, [4, 16], 1), (221, 777, [17, 13], 11),
    (150, 64, [10, 16], 9), (100, 3000, [5, 20], 13), (120, 512, [6, 20], 6),
]


@pytest.mark.parametrize("budget", [0, 1 << 20])  # 1 MiB: one chunk per fused launch
def test_fused_launch_over_many_databases(gpu_ctx, budget):
    """Level 0 of fourteen databases' chunks in fused persistent launches (more segments
    than one launch's table holds, or one chunk per launch under a small budget): every
    reply equals the per-chunk launches' and the plain path's, and the oracle's for the
    first query of each database."""
    import xpir_amd
    rng = np.random.default_rng(5)
    dbs, reqs, expect = [], [], []
    try:
        for si, (n, L, dims, nq) in enumerate(FUSED_DBS):
            orc = Oracle(C2, dims[0])
            prm = xpir_amd.params(C2, dims[0])
            db, data = _setup_db(gpu_ctx, orc, prm, n, L, dims, 300 + si)
            dbs.append(db)
            sk = orc.keygen(60 + si)
            for qi in range(nq):
                index = int(rng.integers(0, n))
                q = np.concatenate(orc.make_query(sk, dims, index, 5000 + 100 * si + qi))
                reqs.append((db, gpu_ctx.alloc(q.nbytes).upload(q), gpu_ctx.alloc(db.reply_words * 8)))
                expect.append((orc, sk, dims, L, data[index * L:(index + 1) * L].tobytes(), q, qi))
        print(f"fused: {len(FUSED_DBS)} databases, {len(reqs)} queries, budget {budget}", flush=True)
        batch = [reqs[i] for i in rng.permutation(len(reqs))]
        gpu_ctx.set_option("group_budget", budget)
        gpu_ctx.reply_batch_dev(batch)
        fused = [rd.download() for _, _, rd in reqs]
        gpu_ctx.set_option("batch_fuse", 0)
        for _, _, rd in reqs:
            rd.upload(np.zeros(rd.nbytes // 8, dtype=np.uint64))
        gpu_ctx.reply_batch_dev(batch)
        for k, (db, qd, rd) in enumerate(reqs):
            orc, sk, dims, L, element, q, qi = expect[k]
            assert np.array_equal(fused[k], rd.download()), k
            assert np.array_equal(fused[k], gpu_ctx.reply(db, q)), k
            if qi == 0:
                words = db.export_words()
                levels, off = [], 0
                for nl in dims:
                    levels.append(q[off:off + nl * 2 * orc.mn])
                    off += nl * 2 * orc.mn
                want, _ = orc.generate_reply(words, db.polys, dims, levels)
                assert np.array_equal(fused[k], want), k
                assert extract_reply(orc, sk, fused[k], dims, L) == element
    finally:
        gpu_ctx.set_option("group_budget", 0)
        gpu_ctx.set_option("batch_fuse", 1)
        for _, qd, rd in reqs:
            qd.free()
            rd.free()
        for db in dbs:
            db.free()


def test_fused_launch_mixes_cryptosystems(gpu_ctx):
    """One fused level-0 launch over databases of three cryptosystems (n = 1024, 2048,
    4096; one, two and three moduli): each segment runs at its own degree and modulus
    count; every reply equals the plain path's and the oracle's."""
    import xpir_amd
    c1 = "LWE:80:1024:60"
    # grids no wider in their second dimension than their first: A is sized from the first
    # (setandgetAbsBitPerCiphertext(n0)), so a wider second level can exceed the noise budget
    specs = [(c1, 1024, 5000, [32, 32], 7), (C2, 160, 1024, [16, 10], 9), ("LWE:100:4096:180", 90, 2500, [10, 10], 6),
             (c1, 400, 9000, [40, 10], 12)]
    rng = np.random.default_rng(11)
    dbs, reqs, expect = [], [], []
    try:
        for si, (desc, n, L, dims, nq) in enumerate(specs):
            orc = Oracle(desc, dims[0])
            prm = xpir_amd.params(desc, dims[0])
            db, data = _setup_db(gpu_ctx, orc, prm, n, L, dims, 700 + si)
            dbs.append(db)
            sk = orc.keygen(70 + si)
            for qi in range(nq):
                index = int(rng.integers(0, n))
                q = np.concatenate(orc.make_query(sk, dims, index, 7000 + 100 * si + qi))
                reqs.append((db, gpu_ctx.alloc(q.nbytes).upload(q), gpu_ctx.alloc(db.reply_words * 8)))
                expect.append((orc, sk, dims, L, data[index * L:(index + 1) * L].tobytes(), q, qi))
        print(f"fused mix: {len(specs)} databases of 3 cryptosystems, {len(reqs)} queries", flush=True)
        gpu_ctx.timing(2)
        gpu_ctx.timing_read_batched(reset=True)
        gpu_ctx.reply_batch_dev([reqs[i] for i in rng.permutation(len(reqs))])
        gpu_ctx.sync()
        _, launches, _ = gpu_ctx.timing_read_batched(reset=True)
        gpu_ctx.timing(False)
        assert launches == 1  # every chunk in one fused launch
        for k, (db, qd, rd) in enumerate(reqs):
            orc, sk, dims, L, element, q, qi = expect[k]
            got = rd.download()
            assert np.array_equal(got, gpu_ctx.reply(db, q)), k
            if qi == 0:
                words = db.export_words()
                levels, off = [], 0
                for nl in dims:
                    levels.append(q[off:off + nl * 2 * orc.mn])
                    off += nl * 2 * orc.mn
                want, _ = orc.generate_reply(words, db.polys, dims, levels)
                assert np.array_equal(got, want), k
            assert extract_reply(orc, sk, got, dims, L) == element
    finally:
        gpu_ctx.timing(False)
        for _, qd, rd in reqs:
            qd.free()
            rd.free()
        for db in dbs:
            db.free()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fused_random_mixes_equal_plain(gpu_ctx, seed):
    """Seeded random mixes: 5-8 databases of random element count and size, d = 1, 2 or 3
    over random grids, 1..80 queries each (above 64, one database gives two chunks in the
    same fused launch), interleaved.  Every reply of the fused launch equals the per-chunk
    launches' and the plain path's."""
    import xpir_amd
    rng = np.random.default_rng(1000 + seed)
    dbs, reqs = [], []
    try:
        for si in range(int(rng.integers(5, 9))):
            desc = C2 if rng.random() < 0.7 else "LWE:80:1024:60"
            d = int(rng.choice([1, 2, 2, 3]))
            if d == 1:
                dims = [int(rng.integers(8, 60))]
            elif d == 2:
                d0 = int(rng.integers(4, 24))
                dims = [d0, int(rng.integers(2, d0 + 1))]
            else:
                d0 = int(rng.integers(4, 10))
                dims = [d0, int(rng.integers(2, d0 + 1)), int(rng.integers(2, 4))]
            n = int(rng.integers(max(1, int(np.prod(dims)) // 2), int(np.prod(dims)) + 1))
            L = int(rng.integers(16, 6000))
            orc = Oracle(desc, dims[0])
            prm = xpir_amd.params(desc, dims[0])
            db, _ = _setup_db(gpu_ctx, orc, prm, n, L, dims, 9000 + 100 * seed + si)
            dbs.append(db)
            sk = orc.keygen(90 + si)
            nq = int(rng.integers(1, 81))
            for qi in range(nq):
                q = np.concatenate(orc.make_query(sk, dims, int(rng.integers(0, n)), 90000 + 1000 * si + qi))
                reqs.append((db, gpu_ctx.alloc(q.nbytes).upload(q), gpu_ctx.alloc(db.reply_words * 8), q))
        print(f"random mix {seed}: {len(dbs)} databases, {len(reqs)} queries", flush=True)
        batch = [reqs[i][:3] for i in rng.permutation(len(reqs))]
        gpu_ctx.reply_batch_dev(batch)
        fused = [rd.download() for _, _, rd, _ in reqs]
        gpu_ctx.set_option("batch_fuse", 0)
        for _, _, rd, _ in reqs:
            rd.upload(np.zeros(rd.nbytes // 8, dtype=np.uint64))
        gpu_ctx.reply_batch_dev(batch)
        for k, (db, qd, rd, q) in enumerate(reqs):
            assert np.array_equal(fused[k], rd.download()), k
            assert np.array_equal(fused[k], gpu_ctx.reply(db, q)), k
    finally:
        gpu_ctx.set_option("batch_fuse", 1)
        for _, qd, rd, _ in reqs:
            qd.free()
            rd.free()
        for db in dbs:
            db.free()
