"""Level 0 on the matrix cores (context option db_layout = 1).  A database imported in the
matrix layout must answer with exactly the words of the same database in the row-major
layout (whose replies the oracle and the reference's golden vectors pin in the other GPU
tests), for single and batched replies, odd query counts, unit counts that are not a
multiple of four, two polys per element, d = 1, 2 and 3, prefixes longer than one int32
pass, and chains whose last steps cannot settle their excess multiple of p (recomputed
whole): words 1 against query words p - 1 keep every step's residue wrapping."""
import numpy as np
import pytest

from oracle.oracle import Oracle, extract_reply

pytestmark = pytest.mark.gpu

C2 = "LWE:80:2048:120"


def _pair(ctx, prm, data, n, L, dims):
    ctx.set_option("db_layout", 0)
    row = ctx.import_db(prm, data, n, L, dims)
    ctx.set_option("db_layout", 1)
    try:
        mx = ctx.import_db(prm, data, n, L, dims)
    finally:
        ctx.set_option("db_layout", 0)
    return row, mx


@pytest.mark.parametrize("n,L,dims,nq", [
    (100 * 80, 256, [100, 80], 5),     # tail of 68 after one block of 32
    (300 * 7, 1024, [300, 7], 9),      # 7 units: the last tile of four half empty
    (200 * 10, 16384, [200, 10], 3),   # two polys per element
    (5000, 64, [5000], 2),             # d = 1
    (96 * 4 * 3, 200, [96, 4, 3], 4),  # d = 3
    (130, 100, [130], 1),              # one query: the single-row-tile kernel
])
def test_matrix_layout_equals_row_major(gpu_ctx, n, L, dims, nq):
    import xpir_amd
    print(f"matrix layout case n={n} L={L} dims={dims} nq={nq}", flush=True)
    orc = Oracle(C2, dims[0])
    prm = xpir_amd.params(C2, dims[0])
    data = orc.rand_bytes(700 + n, 0, n * L)
    row, mx = _pair(gpu_ctx, prm, data, n, L, dims)
    rng = np.random.default_rng(n)
    sk = orc.keygen(77)
    reqs, queries, idx = [], [], []
    for qi in range(nq):
        index = int(rng.integers(0, n))
        q = np.concatenate(orc.make_query(sk, dims, index, 500 + qi))
        qd = gpu_ctx.alloc(q.nbytes).upload(q)
        rd = gpu_ctx.alloc(mx.reply_words * 8)
        reqs.append((mx, qd, rd))
        queries.append(q)
        idx.append(index)
    gpu_ctx.reply_batch_dev(reqs)
    for (db, qd, rd), q, index in zip(reqs, queries, idx):
        want = gpu_ctx.reply(row, q)
        assert np.array_equal(rd.download(), want)
        assert np.array_equal(gpu_ctx.reply(mx, q), want)
        assert extract_reply(orc, sk, want, dims, L) == data[index * L:(index + 1) * L].tobytes()
    for _, qd, rd in reqs:
        qd.free()
        rd.free()
    row.free()
    mx.free()


def test_matrix_layout_matches_oracle(gpu_ctx):
    import xpir_amd
    n, L, dims = 90 * 12, 512, [90, 12]
    orc = Oracle(C2, dims[0])
    prm = xpir_amd.params(C2, dims[0])
    data = orc.rand_bytes(31, 0, n * L)
    row, mx = _pair(gpu_ctx, prm, data, n, L, dims)
    sk = orc.keygen(5)
    q = np.concatenate(orc.make_query(sk, dims, 17, 9))
    words = row.export_words()
    levels, off = [], 0
    for nl in dims:
        levels.append(q[off:off + nl * 2 * orc.mn])
        off += nl * 2 * orc.mn
    want, _ = orc.generate_reply(words, row.polys, dims, levels)
    assert np.array_equal(gpu_ctx.reply(mx, q), want)
    row.free()
    mx.free()


def _ntt_pair(ctx, prm, words, units, P, dims):
    ctx.set_option("db_layout", 0)
    row = ctx.import_ntt_words(prm, words, units, P, 64, dims)
    ctx.set_option("db_layout", 1)
    try:
        mx = ctx.import_ntt_words(prm, words, units, P, 64, dims)
    finally:
        ctx.set_option("db_layout", 0)
    return row, mx


@pytest.mark.parametrize("kind", ["ones", "lazy_max", "random_lazy"])
def test_matrix_layout_adversarial_words(gpu_ctx, kind):
    """Words 1 against query words p - 1: each step adds p - 1, so the residue wraps at
    almost every step and no tail settles its start: every chain is recomputed whole.
    Words 2p - 1 and random words in [0, 2p): the largest lazy words the import makes."""
    import xpir_amd
    dims = [120, 6]
    prm = xpir_amd.params(C2, dims[0])
    orc = Oracle(C2, dims[0])
    mn = orc.mn
    units = dims[0] * dims[1]
    P = 1
    ps = np.array([prm.moduli[c] for c in range(prm.m)], dtype=np.uint64)
    rng = np.random.default_rng(3)
    if kind == "ones":
        words = np.ones(units * P * mn, dtype=np.uint64)
    elif kind == "lazy_max":
        words = np.tile(np.repeat(2 * ps - 1, prm.n), units * P).astype(np.uint64)
    else:
        hi = np.tile(np.repeat(2 * ps, prm.n), units * P)
        words = (rng.integers(0, 2**62, size=units * P * mn, dtype=np.uint64) % hi).astype(np.uint64)
    row, mx = _ntt_pair(gpu_ctx, prm, words, units, P, dims)
    qw = row.query_words
    pm1 = np.tile(np.repeat(ps - 1, prm.n), qw // mn).astype(np.uint64)
    if kind == "ones":
        q = pm1
    else:
        q = (rng.integers(0, 2**62, size=qw, dtype=np.uint64) % (pm1 + 1)).astype(np.uint64)
    want = gpu_ctx.reply(row, q)
    assert np.array_equal(gpu_ctx.reply(mx, q), want)
    row.free()
    mx.free()


def test_matrix_layout_query_words_above_modulus(gpu_ctx):
    """A query word not below its modulus takes the reference's step over the whole chain
    (its lazy products leave [0, 2p)): the words still equal the row-major layout's."""
    import xpir_amd
    n, L, dims = 100 * 8, 128, [100, 8]
    orc = Oracle(C2, dims[0])
    prm = xpir_amd.params(C2, dims[0])
    data = orc.rand_bytes(8, 0, n * L)
    row, mx = _pair(gpu_ctx, prm, data, n, L, dims)
    sk = orc.keygen(6)
    q = np.concatenate(orc.make_query(sk, dims, 3, 4))
    q[5] = np.uint64(prm.moduli[0] + 12345)  # element 0, half a, modulus 0
    q2 = np.concatenate(orc.make_query(sk, dims, 4, 5))
    qd = [gpu_ctx.alloc(x.nbytes).upload(x) for x in (q, q2)]
    rd = [gpu_ctx.alloc(mx.reply_words * 8) for _ in range(2)]
    gpu_ctx.reply_batch_dev([(mx, qd[0], rd[0]), (mx, qd[1], rd[1])])
    assert np.array_equal(rd[0].download(), gpu_ctx.reply(row, q))
    assert np.array_equal(rd[1].download(), gpu_ctx.reply(row, q2))
    for b in qd + rd:
        b.free()
    row.free()
    mx.free()


def test_matrix_layout_long_prefix(gpu_ctx):
    """A d = 1 level 0 of 70 000 elements: the prefix takes two int32 passes."""
    import xpir_amd
    n, L, dims = 70000, 16, [70000]
    orc = Oracle(C2, dims[0])
    prm = xpir_amd.params(C2, dims[0])
    data = orc.rand_bytes(12, 0, n * L)
    row, mx = _pair(gpu_ctx, prm, data, n, L, dims)
    sk = orc.keygen(8)
    q = np.concatenate(orc.make_query(sk, dims, 69999, 2))
    want = gpu_ctx.reply(row, q)
    assert np.array_equal(gpu_ctx.reply(mx, q), want)
    assert extract_reply(orc, sk, want, dims, L) == data[69999 * L:70000 * L].tobytes()
    row.free()
    mx.free()


def test_matrix_layout_entry_points(gpu_ctx, tmp_path):
    """No-allocation and captured replies run the matrix level 0 too; the calls that need
    row-major words refuse the layout with an error."""
    import xpir_amd
    n, L, dims = 80 * 4, 64, [80, 4]
    orc = Oracle(C2, dims[0])
    prm = xpir_amd.params(C2, dims[0])
    data = orc.rand_bytes(9, 0, n * L)
    row, mx = _pair(gpu_ctx, prm, data, n, L, dims)
    sk = orc.keygen(1)
    q = np.concatenate(orc.make_query(sk, dims, 7, 3))
    want = gpu_ctx.reply(row, q)
    qd = gpu_ctx.alloc(q.nbytes).upload(q)
    rd = gpu_ctx.alloc(mx.reply_words * 8)
    gpu_ctx.reply_dev_noalloc(mx, qd, rd)
    assert np.array_equal(rd.download(), want)
    gpu_ctx.set_option("graphs", 1)
    try:
        for _ in range(3):
            gpu_ctx.reply_dev(mx, qd, rd)
            gpu_ctx.sync()
            assert np.array_equal(rd.download(), want)
    finally:
        gpu_ctx.set_option("graphs", 0)
    with pytest.raises(xpir_amd.XpgError):
        mx.export_words()
    with pytest.raises(xpir_amd.XpgError):
        gpu_ctx.save_db(mx, str(tmp_path / "mx.ntt"))
    with pytest.raises(xpir_amd.XpgError):
        gpu_ctx.autotune(mx)
    qd.free()
    rd.free()
    row.free()
    mx.free()
