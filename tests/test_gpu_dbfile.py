"""The NTT database in a file (xpg_db_save / xpg_db_load; no reference counterpart: the
reference imports at every start).  A loaded database holds the imported words and
answers with the same reply bytes, on any context of the GPU; loads refuse a file made
under other parameters or from another database (digest), corrupted words (checksum),
a truncated or extended file, and a file of another format."""
import hashlib
import os
import shutil

import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

C1, C2, C3 = "LWE:80:1024:60", "LWE:80:2048:120", "LWE:100:4096:180"


def _db(ctx, desc, n, L, dims, alpha, seed):
    import xpir_amd
    orc = Oracle(desc, dims[0])
    prm = xpir_amd.params(desc, dims[0])
    data = orc.rand_bytes(seed, 0, n * L)
    db = ctx.import_db(prm, data, n, L, dims, alpha=alpha)
    return orc, prm, data, db


@pytest.mark.parametrize("desc,n,L,dims,alpha,zero_pad", [
    (C1, 100, 2000, [8, 5], 3, 0),      # a partial last aggregate (100 mod 3 = 1): the reference's bytes
    (C1, 100, 2000, [34], 3, 1),        # ... zero-padded instead, d = 1
    (C2, 300, 13000, [20, 15], 1, 0),   # two polys per element
    (C3, 40, 1024, [8, 5], 1, 0),       # three moduli, A > 56
])
def test_loaded_database_answers_as_imported(gpu_ctx, tmp_path, desc, n, L, dims, alpha, zero_pad):
    import xpir_amd
    print(f"dbfile: {desc} {n} x {L} B dims {dims} alpha {alpha} zero_pad {zero_pad}", flush=True)
    gpu_ctx.set_option("alpha_zero_pad", zero_pad)
    try:
        orc, prm, data, db = _db(gpu_ctx, desc, n, L, dims, alpha, 91)
    finally:
        gpu_ctx.set_option("alpha_zero_pad", 0)
    digest = hashlib.sha256(data.tobytes()).digest()
    path = str(tmp_path / "db.ntt")
    gpu_ctx.save_db(db, path, digest)
    words = db.export_words()
    assert os.path.getsize(path) == 512 + words.nbytes
    assert os.listdir(tmp_path) == ["db.ntt"]  # no temporary file left
    sk = orc.keygen(92)
    q = np.concatenate(orc.make_query(sk, dims, 7, 93))
    want = gpu_ctx.reply(db, q)
    with xpir_amd.Context(0) as other:  # any context of the GPU loads it
        db2 = other.load_db(prm, path, digest)
        assert db2.dims == dims and db2.polys == db.polys and db2.raw_bytes == db.raw_bytes
        assert np.array_equal(db2.export_words(), words)
        assert np.array_equal(other.reply(db2, q), want)
        db2.free()
    db3 = gpu_ctx.load_db(prm, path)  # no digest given: not checked
    assert np.array_equal(gpu_ctx.reply(db3, q), want)
    db3.free()
    db.free()


def test_load_refuses_other_files(gpu_ctx, tmp_path):
    import xpir_amd
    from xpir_amd.capi import XpgError
    orc, prm, data, db = _db(gpu_ctx, C1, 60, 1500, [6, 10], 1, 95)
    digest = hashlib.sha256(data.tobytes()).digest()
    path = str(tmp_path / "db.ntt")
    gpu_ctx.save_db(db, path, digest)
    size = os.path.getsize(path)
    db.free()

    def refused(p, prm_, dig, match):
        with pytest.raises(XpgError, match=match):
            gpu_ctx.load_db(prm_, p, dig).free()

    refused(path, xpir_amd.params(C1, 7), digest, "status -1: .*other parameters")   # other n0: other A
    refused(path, xpir_amd.params(C2, 6), digest, "status -1: .*other parameters")   # other cryptosystem
    refused(path, prm, hashlib.sha256(b"another").digest(), "status -1: .*digest")
    # one flipped bit in the words
    bad = str(tmp_path / "flipped.ntt")
    shutil.copy(path, bad)
    with open(bad, "r+b") as f:
        f.seek(512 + 8 * 1234 + 3)
        b = f.read(1)
        f.seek(-1, os.SEEK_CUR)
        f.write(bytes([b[0] ^ 0x10]))
    refused(bad, prm, digest, "status -1: .*checksum")
    # two swapped words (the checksum is keyed by position)
    swapped = str(tmp_path / "swapped.ntt")
    shutil.copy(path, swapped)
    with open(swapped, "r+b") as f:
        f.seek(512 + 8 * 10)
        w10 = f.read(8)
        f.seek(512 + 8 * 11)
        w11 = f.read(8)
        assert w10 != w11
        f.seek(512 + 8 * 10)
        f.write(w11 + w10)
    refused(swapped, prm, digest, "status -1: .*checksum")
    # truncated / extended
    short = str(tmp_path / "short.ntt")
    with open(path, "rb") as src, open(short, "wb") as dst:
        dst.write(src.read(size - 8))
    refused(short, prm, digest, "status -4: ")
    longer = str(tmp_path / "long.ntt")
    shutil.copy(path, longer)
    with open(longer, "ab") as f:
        f.write(b"\0" * 8)
    refused(longer, prm, digest, "status -4: ")
    # not this format; no such file
    junk = str(tmp_path / "junk.ntt")
    with open(junk, "wb") as f:
        f.write(os.urandom(600))
    refused(junk, prm, digest, "status -1: .*not a database file")
    refused(str(tmp_path / "missing.ntt"), prm, digest, "status -1: .*cannot open")
    # the original still loads
    gpu_ctx.load_db(prm, path, digest).free()
