"""Error paths of the C ABI on a real context: malformed arguments come back as status
codes with a message (the reference prints or exit()s; this ABI never aborts), and the
context stays usable afterwards."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C2 = "LWE:80:2048:120"


def _raises(fn, code):
    import xpir_amd
    with pytest.raises(xpir_amd.capi.XpgError) as e:
        fn()
    assert f"status {code}" in str(e.value)


def test_import_argument_errors(gpu_ctx):
    import xpir_amd
    prm = xpir_amd.params(C2, 8)
    data = np.zeros(8 * 64, dtype=np.uint8)
    _raises(lambda: gpu_ctx.import_db(prm, data, 8, 64, [8], alpha=0), -1)        # alpha 0
    _raises(lambda: gpu_ctx.import_db(prm, data, 8, 64, [0]), -1)                 # zero dimension
    _raises(lambda: gpu_ctx.import_db(prm, data, 8, 64, [4]), -1)                 # dims[0] != n0
    _raises(lambda: gpu_ctx.import_db(prm, data, 8, 64, [8, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1]), -1)  # d > 10
    prm7 = xpir_amd.params(C2, 7)
    _raises(lambda: gpu_ctx.import_db(prm7, data, 8, 64, [7]), -1)               # grid too small
    empty = np.zeros(0, dtype=np.uint8)
    _raises(lambda: gpu_ctx.import_db(prm, empty, 0, 64, [8]), -1)                # no elements
    _raises(lambda: gpu_ctx.import_db(prm, empty, 8, 0, [8]), -1)                 # empty elements
    # the context still works
    db = gpu_ctx.import_db(prm, data, 8, 64, [8])
    assert db.polys == 1
    db.free()


def test_reply_argument_errors(gpu_ctx):
    import xpir_amd
    prm = xpir_amd.params(C2, 8)
    data = np.arange(8 * 64, dtype=np.uint8)
    db = gpu_ctx.import_db(prm, data, 8, 64, [8])
    try:
        q = np.zeros(db.query_words - 1, dtype=np.uint64)
        _raises(lambda: gpu_ctx.reply(db, q), -4)                                   # wrong query size
        _raises(lambda: gpu_ctx.set_option("mac_kernel", 9), -1)                   # unknown value
        _raises(lambda: gpu_ctx.set_option("no_such_knob", 1), -1)
        rep = gpu_ctx.reply(db, np.zeros(db.query_words, dtype=np.uint64))          # still usable
        assert rep.size == db.reply_words
    finally:
        db.free()


def test_empty_batch_and_bad_requests(gpu_ctx):
    import xpir_amd
    gpu_ctx.reply_batch_dev([])                                                     # nothing to do
    prm = xpir_amd.params(C2, 8)
    db = gpu_ctx.import_db(prm, np.arange(8 * 64, dtype=np.uint8), 8, 64, [8])
    try:
        lib = xpir_amd.capi.lib()
        req = (xpir_amd.capi.XpgRequest * 1)()
        req[0].db, req[0].query_dev, req[0].reply_dev = db.h, None, None
        assert lib.xpg_reply_batch_dev(gpu_ctx.h, req, 1) == -1                    # null buffers
        assert lib.xpg_autotune(gpu_ctx.h, db.h, 65, None, 0, None) == -1          # batch above 64
        assert lib.xpg_autotune(gpu_ctx.h, None, 0, None, 0, None) == -1
    finally:
        db.free()


def test_database_outlives_its_context():
    """A database stays resident after the context that imported it is destroyed, serves
    replies on another context of the same GPU, and frees cleanly afterwards (the
    reference keeps an imported_database past its session's reply generator)."""
    import xpir_amd
    from oracle.oracle import Oracle
    prm = xpir_amd.params(C2, 8)
    orc = Oracle(C2, 8)
    rng = np.random.default_rng(11)
    data = rng.integers(0, 256, 8 * 3000, dtype=np.uint8)
    q = np.concatenate([rng.integers(0, p, orc.n, dtype=np.uint64)
                        for _ in range(8) for _ in range(2) for p in orc.p])
    first = xpir_amd.Context(0)
    db = first.import_db(prm, data, 8, 3000, [8])
    want = first.reply(db, q)
    first.close()
    with xpir_amd.Context(0) as second:
        got = second.reply(db, q)
        db.free()
        again = second.import_db(prm, data, 8, 3000, [8])
        assert np.array_equal(second.reply(again, q), want)
        again.free()
    assert np.array_equal(got, want)
