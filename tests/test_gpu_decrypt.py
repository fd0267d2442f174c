"""Decryption on the GPU (xpg_decrypt_dev) against the oracle's decrypt (the reference's
NFLLWE::dec -> invnttAndPowInvPhi -> centred CRT lift -> low A bits, restated in C):
same plaintext bytes for random lazy ciphertexts at every parameter shape, and the
level-by-level extraction of real replies returns the chosen element."""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

C1 = "LWE:80:1024:60"
C2 = "LWE:80:2048:120"
C3 = "LWE:100:4096:180"


@pytest.mark.parametrize("desc,n0", [(C1, 1000), (C2, 65536), (C3, 1024), ("LWE:80:1024:120", 64),
                                      ("LWE:80:2048:180", 10)])
def test_decrypt_random_ciphertexts_match_oracle(gpu_ctx, desc, n0):
    import xpir_amd
    orc = Oracle(desc, n0)
    prm = xpir_amd.params(desc, n0)
    rng = np.random.default_rng(n0 + prm.m)
    sk = orc.keygen(77)
    n_ct = 5
    cts = []
    for t in range(n_ct):
        ct = np.concatenate([rng.integers(0, 2 * p, orc.n, dtype=np.uint64)  # lazy words, as replies hold
                             for _ in range(2) for p in orc.p])
        if t == 1:  # canonical words only
            ct = np.concatenate([rng.integers(0, p, orc.n, dtype=np.uint64) for _ in range(2) for p in orc.p])
        if t == 2:  # a fresh encryption of a 1 at slot 0 (noise stays small)
            ct = orc.encrypt_bit(sk, 1, 5, 3)
        cts.append(ct)
    want = b"".join(orc.decrypt(sk, c).tobytes() for c in cts)
    skd = gpu_ctx.alloc(sk.nbytes).upload(sk)
    ctd = gpu_ctx.alloc(n_ct * 16 * orc.mn).upload(np.concatenate(cts))
    out = gpu_ctx.alloc(n_ct * orc.A * orc.n // 8)
    try:
        gpu_ctx.decrypt_dev(prm, skd, ctd.ptr, n_ct, out)
        got = out.download(np.uint8).tobytes()
        assert got == want
    finally:
        for b in (skd, ctd, out):
            b.free()


@pytest.mark.parametrize("desc,n_elems,elem_bytes,dims,index", [
    (C1, 30, 10240, [30], 17),
    (C2, 50, 4096, [10, 5], 33),
    (C3, 24, 1024, [3, 4, 2], 21),
])
def test_gpu_extraction_returns_the_element(gpu_ctx, desc, n_elems, elem_bytes, dims, index):
    import xpir_amd
    orc = Oracle(desc, dims[0])
    prm = xpir_amd.params(desc, dims[0])
    data = orc.rand_bytes(91, 0, n_elems * elem_bytes)
    db = gpu_ctx.import_db(prm, data, n_elems, elem_bytes, dims)
    sk = orc.keygen(92)
    skd = gpu_ctx.alloc(sk.nbytes).upload(sk)
    qd = gpu_ctx.alloc(db.query_words * 8)
    gpu_ctx.query_gen_dev(prm, skd, dims, index, 93, qd)
    rd = gpu_ctx.alloc(db.reply_words * 8)
    try:
        gpu_ctx.reply_dev(db, qd, rd)
        got = gpu_ctx.extract_reply_dev(prm, skd, rd.ptr, dims, elem_bytes)
        assert got == data[index * elem_bytes:(index + 1) * elem_bytes].tobytes()
    finally:
        for b in (skd, qd, rd):
            b.free()
        db.free()


def test_decrypt_rejects_bad_arguments(gpu_ctx):
    import xpir_amd
    prm = xpir_amd.params(C2, 1000)
    buf = gpu_ctx.alloc(1 << 16)
    try:
        bad = xpir_amd.params(C2, 1000)
        bad.A = 200
        with pytest.raises(xpir_amd.capi.XpgError):
            gpu_ctx.decrypt_dev(bad, buf, buf.ptr, 1, buf)
        # zero ciphertexts: nothing to do
        gpu_ctx.decrypt_dev(prm, buf, buf.ptr, 0, buf)
    finally:
        buf.free()
