"""The MAC kernels (one chain per lane, and streaming block summaries with a max-plus
scan) against the oracle, at shapes that exercise partial blocks, several polys per
element, several groups, and the full single-query configuration against the golden
reply bytes written from the reference's own arithmetic."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _random_level(orc, rng, n_l, groups, P):
    db = np.concatenate([rng.integers(0, 2 * p, orc.n, dtype=np.uint64)
                         for _ in range(groups * n_l * P) for p in orc.p])
    q = np.concatenate([rng.integers(0, p, orc.n, dtype=np.uint64)
                        for _ in range(n_l) for _ in range(2) for p in orc.p])
    return db, q


@pytest.mark.parametrize("mode", [1, 3])
@pytest.mark.parametrize("desc,n_l,groups,P", [
    ("LWE:80:2048:120", 1, 1, 1),
    ("LWE:80:2048:120", 63, 1, 1),
    ("LWE:80:2048:120", 1000, 1, 1),
    ("LWE:80:2048:120", 2049, 1, 2),      # tiles of 128 elements, one partial
    ("LWE:80:1024:60", 1000, 1, 6),       # six polys per element
    ("LWE:100:4096:180", 300, 3, 1),      # several groups: level 0 of a 2-level grid
])
def test_mac_kernels_match_oracle(gpu_ctx, mode, desc, n_l, groups, P):
    import xpir_amd
    orc = Oracle(desc, n_l)
    prm = xpir_amd.params(desc, n_l)
    rng = np.random.default_rng(n_l * 7 + groups)
    db_words, q0 = _random_level(orc, rng, n_l, groups, P)
    # adversarial words: DB words at the top of [0,2p) push chains into [p,3p)
    db_words.reshape(-1, orc.m, orc.n)[::5, 0, :] = np.uint64(2 * orc.p[0] - 1)
    dims = [n_l, groups] if groups > 1 else [n_l]
    levels = [q0]
    if groups > 1:
        levels.append(np.concatenate([rng.integers(0, p, orc.n, dtype=np.uint64)
                                      for _ in range(groups) for _ in range(2) for p in orc.p]))
    want, _ = orc.generate_reply(db_words, P, dims, levels)
    db = gpu_ctx.import_ntt_words(prm, db_words, groups * n_l, P, 8, dims)
    gpu_ctx.set_option("mac_kernel", mode)
    try:
        got = gpu_ctx.reply(db, np.concatenate(levels))
        assert np.array_equal(got, want)
    finally:
        gpu_ctx.set_option("mac_kernel", 0)
        db.free()


def test_single_query_full_size_matches_golden(gpu_ctx):
    """65536 x 4 KiB, LWE:80:2048:120: reply bytes == the reference's (golden SHA-256),
    with every MAC kernel choice."""
    import xpir_amd
    from xpir_amd.workloads import C2_SINGLE as cfg
    gold = json.load(open(os.path.join(HERE, "golden", "golden.json")))["c2_single"]
    prm = xpir_amd.params(cfg["crypto"], cfg["dims"][0])
    n, L = cfg["n_elems"], cfg["elem_bytes"]
    raw = gpu_ctx.alloc(n * L + 16)
    gpu_ctx.random_bytes_dev(cfg["db_seed"], 0, raw, n * L)
    db = gpu_ctx.import_db_dev(prm, raw, n, L, cfg["dims"])
    raw.free()
    words = db.export_words()
    assert hashlib.sha256(words.tobytes()).hexdigest() == gold["ntt_sha256"]
    del words
    sk = xpir_amd.keygen(prm, cfg["key_seed"])
    skd = gpu_ctx.alloc(sk.nbytes).upload(sk)
    qd = gpu_ctx.alloc(db.query_words * 8)
    gpu_ctx.query_gen_dev(prm, skd, cfg["dims"], cfg["index"], cfg["noise_seed"], qd)
    rd = gpu_ctx.alloc(db.reply_words * 8)
    try:
        for mode in (3, 1, 0):
            gpu_ctx.set_option("mac_kernel", mode)
            gpu_ctx.reply_dev(db, qd, rd)
            rep = rd.download()
            assert hashlib.sha256(rep.tobytes()).hexdigest() == gold["reply_sha256"], f"mode {mode}"
    finally:
        gpu_ctx.set_option("mac_kernel", 0)
        for b in (skd, qd, rd):
            b.free()
        db.free()


@pytest.mark.parametrize("pf", [3, 1, 2])  # load then compute, double-buffered, rolling
@pytest.mark.parametrize("u,e", [(4, 16), (8, 64), (8, 256), (16, 512), (8, 4096), (16, 48), (4, 1000), (8, 8),
                                 (8, 24), (4, 12)])
def test_streaming_block_shapes_match_oracle(gpu_ctx, u, e, pf):
    """The streaming MAC (block summaries + scan) at several block and chunk sizes: partial
    last blocks and chunks, one block per chain, many blocks per chain."""
    import xpir_amd
    desc, n_l = "LWE:80:2048:120", 3001
    orc = Oracle(desc, n_l)
    prm = xpir_amd.params(desc, n_l)
    rng = np.random.default_rng(123)
    db_words, q = _random_level(orc, rng, n_l, 1, 1)
    db_words.reshape(-1, orc.m, orc.n)[::7, 1, :] = np.uint64(2 * orc.p[1] - 1)
    want = orc.reply_level(db_words, q, n_l, 1, 1)
    db = gpu_ctx.import_ntt_words(prm, db_words, n_l, 1, 8, [n_l])
    gpu_ctx.set_option("mac_kernel", 3)
    gpu_ctx.set_option("stream_u", u)
    gpu_ctx.set_option("stream_e", e)
    gpu_ctx.set_option("stream_pf", pf)
    try:
        assert np.array_equal(gpu_ctx.reply(db, q), want)
    finally:
        for k in ("mac_kernel", "stream_u", "stream_e", "stream_pf"):
            gpu_ctx.set_option(k, 0)
        db.free()
