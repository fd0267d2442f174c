"""The MAC autotuner (xpg_autotune) and its per-device table: every level of a database
and every batch size gets an entry; replies and batched replies after tuning are the
same words as before tuning and as the oracle's; a saved table can be cleared and
installed again; entries naming kernels that are not built are refused."""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

C1 = "LWE:80:1024:60"
C2 = "LWE:80:2048:120"
C3 = "LWE:100:4096:180"


@pytest.mark.parametrize("desc,n_elems,elem_bytes,dims", [
    (C1, 40, 10240, [40]),
    (C2, 300, 256, [20, 15]),
    (C3, 16, 1024, [4, 4]),
    (C2, 2000, 512, [40, 50]),
])
def test_tuned_replies_match_untuned_and_oracle(desc, n_elems, elem_bytes, dims, tmp_path):
    import xpir_amd
    orc = Oracle(desc, dims[0])
    data = orc.rand_bytes(51, 0, n_elems * elem_bytes)
    ctx = xpir_amd.Context(0)
    bufs = []
    try:
        ctx.tune_clear()
        prm = xpir_amd.params(desc, dims[0])
        db = ctx.import_db(prm, data, n_elems, elem_bytes, dims)
        sk = orc.keygen(52)
        skd = ctx.alloc(sk.nbytes).upload(sk)
        bufs.append(skd)
        qs, untuned = [], []
        for k in range(4):
            qd = ctx.alloc(db.query_words * 8)
            ctx.query_gen_dev(prm, skd, dims, (7 * k + 3) % n_elems, 60 + k, qd)
            rd = ctx.alloc(db.reply_words * 8)
            ctx.reply_dev(db, qd, rd)
            untuned.append(rd.download())
            qs.append((qd, rd))
            bufs += [qd, rd]
        if n_elems * elem_bytes <= 1 << 20:
            # the untuned words are the oracle's
            P = orc.polys(elem_bytes)
            words = np.concatenate([orc.import_element(data[e * elem_bytes:(e + 1) * elem_bytes])
                                    for e in range(n_elems)])
            units = int(np.prod(dims))
            full = np.zeros(units * P * orc.mn, dtype=np.uint64)
            full[: words.size] = words
            qh = qs[0][0].download()
            off, ql = 0, []
            for n_l in dims:
                ql.append(qh[off:off + n_l * 2 * orc.mn])
                off += n_l * 2 * orc.mn
            assert np.array_equal(orc.generate_reply(full, P, dims, ql)[0], untuned[0])
        entries = ctx.autotune(db, max_batch=4)
        kinds = [(e["kind"], e["n_l"], e["groups"], e["qn"]) for e in entries]
        assert sum(1 for e in entries if e["kind"] == 0) == len(dims)
        assert {e["qn"] for e in entries if e["kind"] == 1} == {2, 4}, kinds
        for e in entries:
            assert e["ms"] > 0 and e["ms"] <= e["default_ms"] * 1.5, e
        for (qd, rd), want in zip(qs, untuned):
            ctx.reply_dev(db, qd, rd)
            assert np.array_equal(rd.download(), want)
        # a batch after tuning (level 0 in one launch with the tuned register blocks)
        outs = [ctx.alloc(db.reply_words * 8) for _ in qs]
        bufs += outs
        ctx.reply_batch_dev([(db, qd, o) for (qd, _), o in zip(qs, outs)])
        for o, want in zip(outs, untuned):
            assert np.array_equal(o.download(), want)
        # save, clear, reinstall (through a JSON file)
        saved = ctx.tune_get()
        assert len(saved) >= len(entries)
        path = tmp_path / "tune.json"
        ctx.tune_save(str(path))
        ctx.tune_clear()
        assert ctx.tune_get() == []
        ctx.tune_load(str(path))
        assert len(ctx.tune_get()) == len(saved)
        qd, rd = qs[1]
        ctx.reply_dev(db, qd, rd)
        assert np.array_equal(rd.download(), untuned[1])
        db.free()
    finally:
        for b in bufs:
            b.free()
        ctx.tune_clear()
        ctx.close()


def test_tune_set_refuses_unbuilt_shapes():
    import xpir_amd
    ctx = xpir_amd.Context(0)
    try:
        base = {"kind": 0, "n": 2048, "m": 2, "P": 1, "n_l": 100, "groups": 4, "qn": 0,
                "kernel": 3, "a": 8, "b": 0, "c": 0, "ms": 1.0, "default_ms": 1.0}
        ctx.tune_set([base])
        for bad in ({"kernel": 2, "a": 8, "b": 4}, {"kernel": 3, "a": 32}, {"kernel": 7},
                    {"kind": 1, "qn": 4, "kernel": 4, "a": 8, "b": 4}, {"kind": 0, "qn": 3}):
            with pytest.raises(xpir_amd.capi.XpgError):
                ctx.tune_set([dict(base, **bad)])
    finally:
        ctx.tune_clear()
        ctx.close()
