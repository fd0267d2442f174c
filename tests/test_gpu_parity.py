"""Parity of the HIP path (libxpir_amd.so through its C ABI) with the oracle.

Bar: bit-exact words for import (absorb + NTT), query encryption, every reply level and
the recursion step, on the same seeded inputs; plus decrypt(reply) == chosen element.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle, extract_reply

pytestmark = pytest.mark.gpu

C1 = "LWE:80:1024:60"
C2 = "LWE:80:2048:120"
C3 = "LWE:100:4096:180"


def _db_bytes(orc, seed, n_elems, elem_bytes):
    return orc.rand_bytes(seed, 0, n_elems * elem_bytes)


def _oracle_import(orc, data, n_elems, elem_bytes, alpha, units_total, zero_pad=False):
    unit = elem_bytes * alpha
    P = orc.polys(unit)
    rows = Oracle.aggregate_units(data, n_elems, elem_bytes, alpha, zero_pad)
    units = rows.shape[0]
    words = [orc.import_element(r) for r in rows]
    out = np.zeros(units_total * P * orc.mn, dtype=np.uint64)
    if words:
        out[: units * P * orc.mn] = np.concatenate(words)
    return out, P


@pytest.mark.parametrize("desc,n_elems,elem_bytes,alpha,dims,zero_pad", [
    (C1, 20, 10240, 1, [20], 0),
    (C1, 13, 1000, 1, [16], 0),       # ragged element size, padding units
    (C2, 9, 4096, 2, [5], 0),         # alpha aggregation with a short last unit: the reference's bytes
    (C2, 9, 4096, 2, [5], 1),         # ... and with option alpha_zero_pad
    (C2, 2, 4096, 3, [1], 0),         # one partial unit, no unit before it
    (C1, 11, 3000, 4, [3, 2], 0),     # partial last unit of 3, padding units after it
    (C2, 7, 4101, 1, [7], 0),
    (C3, 5, 1024, 1, [6], 0),         # A = 78: two sub-chunks per coefficient
    (C3, 5, 1024, 1, [1024], 0),      # n0 = 1024: A = 74, the recursive configuration's
    (C3, 3, 40000, 1, [3], 0),
])
def test_import_matches_oracle(gpu_ctx, desc, n_elems, elem_bytes, alpha, dims, zero_pad):
    import xpir_amd
    orc = Oracle(desc, dims[0])
    prm = xpir_amd.params(desc, dims[0])
    assert prm.A == orc.A
    if dims == [1024]:
        assert prm.A == 74
    data = _db_bytes(orc, 11, n_elems, elem_bytes)
    gpu_ctx.set_option("alpha_zero_pad", zero_pad)
    try:
        db = gpu_ctx.import_db(prm, data, n_elems, elem_bytes, dims, alpha=alpha)
    finally:
        gpu_ctx.set_option("alpha_zero_pad", 0)
    want, P = _oracle_import(orc, data, n_elems, elem_bytes, alpha, int(np.prod(dims)), bool(zero_pad))
    assert db.polys == P
    got = db.export_words()
    assert np.array_equal(got, want)
    db.free()


def test_alpha_aggregation_matches_reference_golden(gpu_ctx):
    """n_elems not a multiple of alpha: the GPU import's words equal the reference's
    (readAggregatedStream in importDataNFL's loop, then deserializeDataNFL; golden
    hashes), and a reply over them equals the oracle's."""
    import hashlib
    import json
    import os
    import xpir_amd
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))["alpha"]
    for g in gold:
        orc = Oracle(g["crypto"], g["n0"])
        prm = xpir_amd.params(g["crypto"], g["n0"])
        n, L, alpha, units = g["n_elems"], g["elem_bytes"], g["alpha"], g["units"]
        dims = [g["n0"]] if g["n0"] >= units else [g["n0"], -(-units // g["n0"])]
        data = orc.rand_bytes(g["seed"], 0, n * L)
        db = gpu_ctx.import_db(prm, data, n, L, dims, alpha=alpha)
        try:
            words = db.export_words()
            per = db.polys * orc.mn
            assert db.polys == g["polys"]
            assert hashlib.sha256(words[: units * per].tobytes()).hexdigest() == g["ntt_sha256"], g
            assert not words[units * per:].any()
            sk = orc.keygen(5)
            q = orc.make_query(sk, dims, units - 1, 6)
            want, _ = orc.generate_reply(words, db.polys, dims, q)
            assert np.array_equal(gpu_ctx.reply(db, np.concatenate(q)), want)
        finally:
            db.free()


@pytest.mark.parametrize("desc,dims", [(C1, [7]), (C2, [3, 4]), (C3, [2, 3])])
def test_query_generation_matches_oracle(gpu_ctx, desc, dims):
    import xpir_amd
    orc = Oracle(desc, dims[0])
    prm = xpir_amd.params(desc, dims[0])
    sk = orc.keygen(4)
    assert np.array_equal(sk, xpir_amd.keygen(prm, 4))
    index = int(np.prod(dims)) - 2
    want = np.concatenate(orc.make_query(sk, dims, index, 3))
    skd = gpu_ctx.alloc(sk.nbytes).upload(sk)
    qd = gpu_ctx.alloc(want.nbytes)
    gpu_ctx.query_gen_dev(prm, skd, dims, index, 3, qd)
    got = qd.download()
    assert np.array_equal(got, want)
    skd.free()
    qd.free()


@pytest.mark.parametrize("desc,n_elems,elem_bytes,dims", [
    (C1, 40, 10240, [40]),
    (C1, 37, 3000, [10, 4]),
    (C2, 64, 4096, [64]),
    (C2, 30, 4096, [6, 5]),
    (C3, 16, 1024, [4, 4]),
    (C2, 25, 700, [5, 5, 1]),
    (C2, 50, 512, [5, 11]),           # 11 groups: a full and a partial block of 8
    (C1, 70, 900, [3, 9, 3]),
])
def test_reply_matches_oracle(gpu_ctx, desc, n_elems, elem_bytes, dims):
    import xpir_amd
    orc = Oracle(desc, dims[0])
    prm = xpir_amd.params(desc, dims[0])
    data = _db_bytes(orc, 1, n_elems, elem_bytes)
    units = int(np.prod(dims))
    words, P = _oracle_import(orc, data, n_elems, elem_bytes, 1, units)
    sk = orc.keygen(4)
    index = n_elems - 1
    qlevels = orc.make_query(sk, dims, index, 3)
    want, Pd = orc.generate_reply(words, P, dims, qlevels)
    db = gpu_ctx.import_db(prm, data, n_elems, elem_bytes, dims)
    got = gpu_ctx.reply(db, np.concatenate(qlevels))
    assert db.reply_polys == Pd
    assert np.array_equal(got, want)
    assert extract_reply(orc, sk, got, dims, elem_bytes) == data[index * elem_bytes:(index + 1) * elem_bytes].tobytes()
    db.free()


def test_c1_plain_full(gpu_ctx):
    """LWE:80:1024:60, 1000 x 10 KiB from seed 1, index from seed 2, noise seed 3."""
    import xpir_amd
    from xpir_amd.workloads import C1_PLAIN as cfg
    orc = Oracle(cfg["crypto"], cfg["dims"][0])
    prm = xpir_amd.params(cfg["crypto"], cfg["dims"][0])
    n, L = cfg["n_elems"], cfg["elem_bytes"]
    data = _db_bytes(orc, cfg["db_seed"], n, L)
    index = cfg["index"]
    sk = orc.keygen(cfg["key_seed"])
    q = orc.make_query(sk, cfg["dims"], index, cfg["noise_seed"])
    db = gpu_ctx.import_db(prm, data, n, L, cfg["dims"])
    words, P = _oracle_import(orc, data, n, L, 1, n)
    assert np.array_equal(db.export_words(), words)
    want, _ = orc.generate_reply(words, P, cfg["dims"], q)
    got = gpu_ctx.reply(db, np.concatenate(q))
    assert np.array_equal(got, want)
    assert extract_reply(orc, sk, got, cfg["dims"], L) == data[index * L:(index + 1) * L].tobytes()
    db.free()


def test_recursive_scaled_matches_golden(gpu_ctx):
    """LWE:100:4096:180, d=2 over a 64x64 grid of 1 KiB elements: the level-1 reply and
    the final reply equal the reference's (golden hashes); the reply decrypts."""
    import hashlib
    import json
    import os
    import xpir_amd
    from xpir_amd.workloads import C3_SCALED as cfg
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))["c3_scaled"]
    orc = Oracle(cfg["crypto"], cfg["dims"][0])
    prm = xpir_amd.params(cfg["crypto"], cfg["dims"][0])
    n, L = cfg["n_elems"], cfg["elem_bytes"]
    raw = gpu_ctx.alloc(n * L + 16)
    gpu_ctx.random_bytes_dev(cfg["db_seed"], 0, raw, n * L)
    db = gpu_ctx.import_db_dev(prm, raw, n, L, cfg["dims"])
    raw.free()
    assert hashlib.sha256(db.export_words().tobytes()).hexdigest() == gold["ntt_sha256"]
    sk = orc.keygen(cfg["key_seed"])
    q = np.concatenate(orc.make_query(sk, cfg["dims"], cfg["index"], cfg["noise_seed"]))
    rep = gpu_ctx.reply(db, q)
    assert hashlib.sha256(rep.tobytes()).hexdigest() == gold["reply_sha256"]
    data = orc.rand_bytes(cfg["db_seed"], cfg["index"] * L, L)
    assert extract_reply(orc, sk, rep, cfg["dims"], L) == data.tobytes()
    db.free()


@pytest.mark.parametrize("desc,n_elems,elem_bytes,dims", [
    ("LWE:80:512:60", 20, 700, [20]),          # smallest ring the reference allows (kMinPolyDegree)
    ("LWE:80:512:60", 24, 300, [6, 4]),
    ("LWE:128:8192:240", 6, 20000, [6]),       # largest ring and four moduli (A = 105: two sub-chunks)
    ("LWE:128:8192:240", 8, 5000, [4, 2]),
])
def test_extreme_ring_sizes_match_oracle(gpu_ctx, desc, n_elems, elem_bytes, dims):
    """n = 512 (1 modulus) and n = 8192 (4 moduli): import words, reply words and the
    decrypted element, against the oracle."""
    import xpir_amd
    orc = Oracle(desc, dims[0])
    prm = xpir_amd.params(desc, dims[0])
    assert (prm.n, prm.m, prm.A) == (orc.n, orc.m, orc.A)
    data = _db_bytes(orc, 7, n_elems, elem_bytes)
    units = int(np.prod(dims))
    words, P = _oracle_import(orc, data, n_elems, elem_bytes, 1, units)
    db = gpu_ctx.import_db(prm, data, n_elems, elem_bytes, dims)
    try:
        assert db.polys == P
        assert np.array_equal(db.export_words(), words)
        sk = orc.keygen(8)
        index = n_elems - 2
        qlevels = orc.make_query(sk, dims, index, 9)
        want, Pd = orc.generate_reply(words, P, dims, qlevels)
        got = gpu_ctx.reply(db, np.concatenate(qlevels))
        assert db.reply_polys == Pd
        assert np.array_equal(got, want)
        assert extract_reply(orc, sk, got, dims, elem_bytes) == \
            data[index * elem_bytes:(index + 1) * elem_bytes].tobytes()
    finally:
        db.free()


def test_host_buffers_over_64mib_stage_through_pinned_chunks(gpu_ctx):
    """A query above the 64 MiB staging chunk (70 MiB: 1100 elements of LWE:80:2048:120)
    reaches the device through the two pinned chunks of the host-buffer entry points: the
    host-buffer reply equals the device-buffer reply, a staged upload reads back
    unchanged, and the reply decrypts to the chosen element."""
    import xpir_amd
    n, L, index = 1100, 64, 517
    orc = Oracle(C2, n)
    prm = xpir_amd.params(C2, n)
    data = orc.rand_bytes(91, 0, n * L)
    db = gpu_ctx.import_db(prm, data, n, L, [n])
    sk = xpir_amd.keygen(prm, 92)
    skd = gpu_ctx.alloc(sk.nbytes).upload(sk)
    qd = gpu_ctx.alloc(db.query_words * 8)
    rd = gpu_ctx.alloc(db.reply_words * 8)
    try:
        gpu_ctx.query_gen_dev(prm, skd, [n], index, 93, qd)
        q = qd.download()
        assert q.nbytes > 64 << 20
        gpu_ctx.reply_dev(db, qd, rd)
        want = rd.download()
        assert np.array_equal(gpu_ctx.reply(db, q), want)
        rev = np.ascontiguousarray(q[::-1])
        q2 = gpu_ctx.alloc(rev.nbytes).upload(rev)
        assert np.array_equal(q2.download(), rev)
        q2.free()
        assert gpu_ctx.extract_reply_dev(prm, skd, rd.ptr, [n], L) == data[index * L:(index + 1) * L].tobytes()
    finally:
        for b in (skd, qd, rd):
            b.free()
        db.free()
