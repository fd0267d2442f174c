"""Replies replayed from captured HIP graphs (option "graphs") give the same words as
eager replies and the oracle: on every replay, with other GPU work and host syncs
between replays, with new query contents in the same buffer, after an option change
(re-capture) and after the database is freed and another imported in its place."""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

C1 = "LWE:80:1024:60"
C2 = "LWE:80:2048:120"
C3 = "LWE:100:4096:180"


def _setup(desc, n_elems, elem_bytes, dims, seed):
    orc = Oracle(desc, dims[0])
    data = orc.rand_bytes(seed, 0, n_elems * elem_bytes)
    unit = elem_bytes
    P = orc.polys(unit)
    units = int(np.prod(dims))
    words = np.zeros(units * P * orc.mn, dtype=np.uint64)
    words[: n_elems * P * orc.mn] = np.concatenate(
        [orc.import_element(data[e * unit:(e + 1) * unit]) for e in range(n_elems)])
    sk = orc.keygen(seed + 1)
    qs, wants = [], []
    for k, index in enumerate((n_elems - 1, n_elems // 3)):
        q = orc.make_query(sk, dims, index, seed + 2 + k)
        qs.append(np.concatenate(q))
        wants.append(orc.generate_reply(words, P, dims, q)[0])
    return orc, data, qs, wants


@pytest.mark.parametrize("desc,n_elems,elem_bytes,dims", [
    (C1, 40, 10240, [40]),
    (C2, 30, 4096, [6, 5]),
    (C3, 16, 1024, [4, 4]),
    (C1, 70, 900, [3, 9, 3]),
])
def test_graph_replays_match_oracle(desc, n_elems, elem_bytes, dims):
    import xpir_amd
    orc, data, qs, wants = _setup(desc, n_elems, elem_bytes, dims, 21)
    ctx = xpir_amd.Context(0)
    try:
        ctx.set_option("graphs", 1)
        prm = xpir_amd.params(desc, dims[0])
        db = ctx.import_db(prm, data, n_elems, elem_bytes, dims)
        qd = ctx.alloc(qs[0].nbytes)
        rd = ctx.alloc(wants[0].nbytes)
        scratch = ctx.alloc(1 << 20)
        for i in range(6):
            k = i % 2
            qd.upload(qs[k])                      # new query contents, same buffer
            ctx.reply_dev(db, qd, rd)
            ctx.random_bytes_dev(100 + i, 0, scratch, 1 << 20)  # eager work between replays
            ctx.sync()
            assert np.array_equal(rd.download(), wants[k]), f"call {i}"
        captures, replays = ctx.graph_stats()
        assert captures == 1 and replays == 5
        # an option change re-captures; the words stay the same
        ctx.set_option("mac_kernel", 1)
        for i in range(3):
            qd.upload(qs[i % 2])
            ctx.reply_dev(db, qd, rd)
            assert np.array_equal(rd.download(), wants[i % 2]), f"after option change, call {i}"
        assert ctx.graph_stats()[0] == 2
        # host-buffer replies go through the context's own staging buffers
        for i in range(3):
            assert np.array_equal(ctx.reply(db, qs[i % 2]), wants[i % 2])
        db.free()
        for b in (qd, rd, scratch):
            b.free()
    finally:
        ctx.close()


def test_graph_not_reused_for_a_new_database():
    """Free a database and import another of the same shape (possibly at the same
    address): the next reply with the same buffers must read the new database."""
    import xpir_amd
    desc, n_elems, elem_bytes, dims = C2, 30, 4096, [6, 5]
    ctx = xpir_amd.Context(0)
    try:
        ctx.set_option("graphs", 1)
        prm = xpir_amd.params(desc, dims[0])
        _, data_a, qs_a, wants_a = _setup(desc, n_elems, elem_bytes, dims, 31)
        _, data_b, _, _ = _setup(desc, n_elems, elem_bytes, dims, 41)
        orc = Oracle(desc, dims[0])
        P = orc.polys(elem_bytes)
        words_b = np.concatenate([orc.import_element(data_b[e * elem_bytes:(e + 1) * elem_bytes])
                                  for e in range(n_elems)])
        q = [qs_a[0][:dims[0] * 2 * orc.mn], qs_a[0][dims[0] * 2 * orc.mn:]]
        want_b = orc.generate_reply(words_b, P, dims, q)[0]
        qd = ctx.alloc(qs_a[0].nbytes).upload(qs_a[0])
        rd = ctx.alloc(wants_a[0].nbytes)
        db = ctx.import_db(prm, data_a, n_elems, elem_bytes, dims)
        for _ in range(3):
            ctx.reply_dev(db, qd, rd)
            assert np.array_equal(rd.download(), wants_a[0])
        db.free()
        db = ctx.import_db(prm, data_b, n_elems, elem_bytes, dims)
        for _ in range(3):
            ctx.reply_dev(db, qd, rd)
            assert np.array_equal(rd.download(), want_b)
        db.free()
        qd.free()
        rd.free()
    finally:
        ctx.close()
