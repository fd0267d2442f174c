"""xpg_reply_dev_noalloc (the reply without any allocation, scratch from the library's
static device array) against xpg_reply_dev, through the C ABI.

Bar: bit-exact words.  Two calls in a row, the second over a smaller database (the
array keeps the first call's contents), plus the size limit and the missing-tables
error, both refused before anything is enqueued.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _query(ctx, prm, dims, index, seed):
    import xpir_amd
    sk = xpir_amd.keygen(prm, seed)
    skd = ctx.alloc(sk.nbytes).upload(sk)
    n_q = sum(dims) * 2 * prm.m * prm.n
    qd = ctx.alloc(n_q * 8)
    ctx.query_gen_dev(prm, skd, dims, index, seed + 1, qd)
    skd.free()
    return qd


def _device_db(ctx, prm, seed, n, L, dims):
    raw = ctx.alloc(n * L + 16)
    ctx.random_bytes_dev(seed, 0, raw, n * L)
    db = ctx.import_db_dev(prm, raw, n, L, dims)
    raw.free()
    return db


def test_noalloc_matches_reply_dev_two_calls():
    """A single-query database (streaming MAC: block summaries in the array), then a
    smaller d = 2 database (level buffers in the array, recursion); each equals
    xpg_reply_dev's words, and the small one also the oracle's."""
    import xpir_amd
    print("noalloc: 8192 x 4 KiB LWE:80:2048:120 d=1, then 100 x 2000 B LWE:80:1024:60 d=2 [16, 8]", flush=True)
    with xpir_amd.Context(0) as ctx:
        big_prm = xpir_amd.params("LWE:80:2048:120", 8192)
        big = _device_db(ctx, big_prm, 21, 8192, 4096, [8192])
        big_q = _query(ctx, big_prm, [8192], 4321, 5)
        desc, dims, n, L = "LWE:80:1024:60", [16, 8], 100, 2000
        orc = Oracle(desc, dims[0])
        small_prm = xpir_amd.params(desc, dims[0])
        data = orc.rand_bytes(22, 0, n * L)
        small = ctx.import_db(small_prm, data, n, L, dims)
        small_q = _query(ctx, small_prm, dims, 77, 9)
        r_big, r_small = ctx.alloc(big.reply_words * 8), ctx.alloc(small.reply_words * 8)
        ref_big, ref_small = ctx.alloc(big.reply_words * 8), ctx.alloc(small.reply_words * 8)
        try:
            # the two no-allocation calls back to back, the second smaller
            ctx.reply_dev_noalloc(big, big_q, r_big)
            ctx.reply_dev_noalloc(small, small_q, r_small)
            ctx.sync()
            got_big, got_small = r_big.download(), r_small.download()
            ctx.reply_dev(big, big_q, ref_big)
            ctx.reply_dev(small, small_q, ref_small)
            assert np.array_equal(got_big, ref_big.download())
            want_small = ref_small.download()
            assert np.array_equal(got_small, want_small)
            # and the small reply against the oracle's words
            P = orc.polys(L)
            words = np.zeros(int(np.prod(dims)) * P * orc.mn, dtype=np.uint64)
            words[: n * P * orc.mn] = np.concatenate([orc.import_element(data[e * L:(e + 1) * L])
                                                       for e in range(n)])
            q_host = small_q.download(count=small.query_words)
            q_levels = [q_host[: dims[0] * 2 * orc.mn], q_host[dims[0] * 2 * orc.mn:]]
            want, _ = orc.generate_reply(words, P, dims, q_levels)
            assert np.array_equal(got_small, want)
            # once more at the large size after the small call: same words
            ctx.reply_dev_noalloc(big, big_q, r_big)
            ctx.sync()
            assert np.array_equal(r_big.download(), got_big)
        finally:
            for b in (r_big, r_small, ref_big, ref_small, big_q, small_q):
                b.free()
            big.free()
            small.free()


def test_noalloc_refuses_oversize_before_enqueueing():
    """A d = 2 grid whose level-0 reply (4096 groups x 64 KiB = 256 MiB) exceeds the
    array: XPG_ESIZE, and the reply buffer is untouched."""
    import xpir_amd
    print("noalloc: oversize LWE:80:2048:120 [4, 4096] x 256 B", flush=True)
    with xpir_amd.Context(0) as ctx:
        prm = xpir_amd.params("LWE:80:2048:120", 4)
        db = _device_db(ctx, prm, 23, 4 * 4096, 256, [4, 4096])
        q = _query(ctx, prm, [4, 4096], 5, 3)
        r = ctx.alloc(db.reply_words * 8)
        sentinel = np.full(db.reply_words, 0x5A5A5A5A5A5A5A5A, dtype=np.uint64)
        r.upload(sentinel)
        try:
            with pytest.raises(xpir_amd.capi.XpgError, match="status -4"):
                ctx.reply_dev_noalloc(db, q, r)
            ctx.sync()
            assert np.array_equal(r.download(), sentinel)
        finally:
            for b in (q, r):
                b.free()
            db.free()


def test_noalloc_needs_the_contexts_tables_for_recursion():
    """d = 2 on a context that holds no NTT tables yet: XPG_EINVAL (building them would
    allocate); after an ordinary reply on that context, the same call succeeds."""
    import xpir_amd
    print("noalloc: tables missing on a second context, LWE:80:1024:60 [8, 4]", flush=True)
    desc, dims, n, L = "LWE:80:1024:60", [8, 4], 30, 1500
    orc = Oracle(desc, dims[0])
    prm = xpir_amd.params(desc, dims[0])
    data = orc.rand_bytes(24, 0, n * L)
    with xpir_amd.Context(0) as owner, xpir_amd.Context(0) as other:
        db = owner.import_db(prm, data, n, L, dims)
        q = _query(owner, prm, dims, 11, 13)  # encrypting on `other` would build its tables
        r1, r2 = other.alloc(db.reply_words * 8), other.alloc(db.reply_words * 8)
        try:
            with pytest.raises(xpir_amd.capi.XpgError, match="status -1"):
                other.reply_dev_noalloc(db, q, r1)
            other.reply_dev(db, q, r2)
            other.reply_dev_noalloc(db, q, r1)
            other.sync()
            assert np.array_equal(r1.download(), r2.download())
        finally:
            for b in (q, r1, r2):
                b.free()
            db.free()
