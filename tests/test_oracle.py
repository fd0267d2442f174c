"""CPU checks that pin the oracle (oracle/xpir_oracle.c) before it is trusted.

Each function of the oracle is checked against something independent of its own code:
  * constants: primality, modulus form, root orders (closed forms, Python big ints);
  * twiddle / phi tables: powers recomputed with pow();
  * absorb: direct little-endian bit extraction with Python big ints;
  * NTT: residues against the negacyclic DFT by brute force, and inverse(forward) = id;
  * MAC: residues against sum(x*y) mod p, lazy words bounded by the chain invariant;
  * end to end: decrypt(reply) == chosen element;
  * byte-exact words (lazy representatives, chain order, buffer straddles): the golden
    fixtures written from the reference's own NFLlib (tests/golden/make_golden.py), and
    that build itself when oracle/_ref is present.
A planted mistake (tests/test_oracle.py::test_planted_mistakes_are_caught) shows that
these checks fail on the plausible slips: a dropped lazy subtraction, a transposed
twiddle index, a wrong chain order, a mis-aligned bit window.
"""
import hashlib
import json
import os
import random

import numpy as np
import pytest

from oracle.oracle import Oracle, RefNFL, extract_reply

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))

C1, C2, C3 = "LWE:80:1024:60", "LWE:80:2048:120", "LWE:100:4096:180"


def is_prime(n):
    if n < 2:
        return False
    for sp in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % sp == 0:
            return n == sp
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def bitrev(v, bits):
    return int(format(v, f"0{bits}b")[::-1], 2)


def test_constants():
    L = Oracle.lib()
    for i in range(17):
        p, r = L.xo_modulus(i), L.xo_root(i)
        assert is_prime(p)
        c = (2**61 + 1 - p) >> 14
        assert p == 2**61 - c * 2**14 + 1 and 0 < c < 512
        assert pow(r, 2**14, p) == 1 and pow(r, 2**13, p) == p - 1  # primitive 2^14-th root
        assert L.xo_inv8192(i) * 8192 % p == 1


@pytest.mark.parametrize("desc,n0,A", [(C1, 1000, 15), (C2, 65536, 42), (C3, 1024, 74), (C3, 64, 76)])
def test_params_and_absorption_bits(desc, n0, A):
    import math
    o = Oracle(desc, n0)
    k, n, bits = (int(x) for x in desc.split(":")[1:])
    berr = 5 * k // 2
    assert o.A == A == math.floor(((bits - 1) - math.log2(n0) - math.log2(berr) - math.log2(n)) / 2)
    assert o.m == bits // 60


@pytest.mark.parametrize("desc", [C1, C2, C3])
def test_tables(desc):
    o = Oracle(desc, 1000)
    n = o.n
    phi, w = o.tables("phi"), o.tables("w")
    for c, p in enumerate(o.p):
        f = int(phi[c * n + 1])
        assert pow(f, n, p) == p - 1 and pow(f, 2 * n, p) == 1
        assert all(int(phi[c * n + j]) == pow(f, j, p) for j in (0, 2, 77, n - 1))
        om = f * f % p
        off, size, L = 0, n, 0
        while size >= 2:
            for i in {0, min(1, size // 2 - 1), size // 2 - 1}:
                assert int(w[c * n + off + i]) == pow(om, i * 2**L, p)
            off += size // 2
            size //= 2
            L += 1


@pytest.mark.parametrize("desc,n0,nbytes", [(C1, 1000, 10240), (C2, 65536, 4101), (C3, 1024, 1024),
                                            (C3, 1024, 40000)])
def test_absorb_is_little_endian_bit_extraction(desc, n0, nbytes):
    o = Oracle(desc, n0)
    data = o.rand_bytes(5, 0, nbytes)
    big = int.from_bytes(data.tobytes(), "little")
    out = o.absorb(data).reshape(-1, o.m, o.n)
    A = o.A
    nchunks = -(-nbytes * 8 // A)
    rng = random.Random(1)
    for j in [0, 1, nchunks - 1, nchunks] + [rng.randrange(nchunks) for _ in range(200)]:
        chunk = (big >> (j * A)) & ((1 << A) - 1)
        poly, slot = divmod(j, o.n)
        if poly >= out.shape[0]:
            continue
        for c, p in enumerate(o.p):
            assert int(out[poly, c, slot]) % p == chunk % p
            if A <= 56:
                assert int(out[poly, c, slot]) == chunk


@pytest.mark.parametrize("desc", [C1, C2, C3])
def test_ntt_residues_match_negacyclic_dft(desc):
    o = Oracle(desc, 1000)
    n, logn = o.n, o.n.bit_length() - 1
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.integers(0, p, n, dtype=np.uint64) for p in o.p])
    y = o.ntt_powphi(x)
    phi = o.tables("phi")
    for c, p in enumerate(o.p):
        f = int(phi[c * n + 1])
        xs = [int(v) for v in x[c * n:(c + 1) * n]]
        for pos in list(range(4)) + [n - 1] + list(rng.integers(0, n, 12)):
            pos = int(pos)
            j = bitrev(pos, logn)
            e = 2 * j + 1
            want = sum(xi * pow(f, (i * e) % (2 * n), p) for i, xi in enumerate(xs)) % p
            got = int(y[c * n + pos])
            assert got < 2 * p and got % p == want
    assert np.array_equal(o.invntt(y), x)


def test_mac_residues_and_bounds():
    o = Oracle(C2, 64)
    rng = np.random.default_rng(9)
    nl, P = 64, 1
    db = np.concatenate([rng.integers(0, 2 * p, o.n, dtype=np.uint64) for _ in range(nl) for p in o.p])
    q = np.concatenate([rng.integers(0, p, o.n, dtype=np.uint64) for _ in range(nl) for _ in range(2) for p in o.p])
    out = o.reply_level(db, q, nl, 1, P).reshape(2, o.m, o.n)
    dbv = db.reshape(nl, o.m, o.n)
    qv = q.reshape(nl, 2, o.m, o.n)
    for h in range(2):
        for c, p in enumerate(o.p):
            for s in (0, 5, o.n - 1):
                want = sum(int(dbv[e, c, s]) * int(qv[e, h, c, s]) for e in range(nl)) % p
                got = int(out[h, c, s])
                assert got % p == want and got < 3 * p


def test_chain_parallel_reply_matches_golden():
    """The all-core split (threads over slot blocks) gives the golden C1 reply and the same
    words as the per-poly split on a grid with several groups, polys and moduli."""
    g = GOLD["c1_plain"]
    o = Oracle(g["crypto"], g["dims"][0])
    n, L = g["n_elems"], g["elem_bytes"]
    data = o.rand_bytes(g["db_seed"], 0, n * L)
    words = np.concatenate([o.import_element(data[e * L:(e + 1) * L]) for e in range(n)])
    sk = o.keygen(g["key_seed"])
    q = o.make_query(sk, g["dims"], g["index"], g["noise_seed"])
    rep = o.reply_level_chains(words, q[0], n, 1, g["P0"], 4)
    assert np.array_equal(rep, np.load(os.path.join(HERE, "golden", "c1_plain_reply.npy")))
    o3 = Oracle(C3, 16)
    rng = np.random.default_rng(12)
    nl, groups, P = 16, 3, 2
    db = np.concatenate([rng.integers(0, 2 * p, o3.n, dtype=np.uint64)
                         for _ in range(groups * nl * P) for p in o3.p])
    qq = np.concatenate([rng.integers(0, p, o3.n, dtype=np.uint64) for _ in range(nl) for _ in range(2) for p in o3.p])
    assert np.array_equal(o3.reply_level_chains(db, qq, nl, groups, P, 3), o3.reply_level(db, qq, nl, groups, P))


def test_golden_c1_plain_reply_bytes():
    g = GOLD["c1_plain"]
    o = Oracle(g["crypto"], g["dims"][0])
    n, L = g["n_elems"], g["elem_bytes"]
    data = o.rand_bytes(g["db_seed"], 0, n * L)
    words = np.concatenate([o.import_element(data[e * L:(e + 1) * L]) for e in range(n)])
    assert hashlib.sha256(words.tobytes()).hexdigest() == g["ntt_sha256"]
    sk = o.keygen(g["key_seed"])
    q = o.make_query(sk, g["dims"], g["index"], g["noise_seed"])
    rep, Pd = o.generate_reply(words, g["P0"], g["dims"], q)
    assert Pd == g["Pd"]
    assert hashlib.sha256(rep.tobytes()).hexdigest() == g["reply_sha256"]
    assert np.array_equal(rep, np.load(os.path.join(HERE, "golden", "c1_plain_reply.npy")))
    assert extract_reply(o, sk, rep, g["dims"], L) == data[g["index"] * L:(g["index"] + 1) * L].tobytes()


def test_golden_recursive_scaled_reply_bytes():
    g = GOLD["c3_scaled"]
    o = Oracle(g["crypto"], g["dims"][0])
    n, L = g["n_elems"], g["elem_bytes"]
    data = o.rand_bytes(g["db_seed"], 0, n * L)
    words = np.concatenate([o.import_element(data[e * L:(e + 1) * L]) for e in range(n)])
    sk = o.keygen(g["key_seed"])
    q = o.make_query(sk, g["dims"], g["index"], g["noise_seed"])
    rep0 = o.reply_level(words, q[0], g["dims"][0], g["dims"][1], g["P0"])
    assert hashlib.sha256(rep0.tobytes()).hexdigest() == g["level0_reply_sha256"]
    rep, Pd = o.generate_reply(words, g["P0"], g["dims"], q)
    assert Pd == g["Pd"] == 6
    assert hashlib.sha256(rep.tobytes()).hexdigest() == g["reply_sha256"]
    assert extract_reply(o, sk, rep, g["dims"], L) == data[g["index"] * L:(g["index"] + 1) * L].tobytes()


@pytest.mark.parametrize("vec", GOLD["absorb"], ids=lambda v: f"{v['crypto']}-{v['nbufs']}x{v['bytes_per_buf']}")
def test_golden_absorb_vectors(vec):
    """Chunks straddling buffer boundaries (the recursion layout) and A > 56."""
    o = Oracle(vec["crypto"], vec["n0"])
    data = o.rand_bytes(vec["seed"], 0, vec["nbufs"] * vec["bytes_per_buf"])
    words = o.import_element(data)
    assert words.size // o.mn == vec["polys"]
    assert hashlib.sha256(words.tobytes()).hexdigest() == vec["sha256"]


@pytest.mark.skipif(not RefNFL.available(), reason="reference build (oracle/_ref) absent")
@pytest.mark.parametrize("desc,n0", [(C1, 1000), (C2, 65536), (C3, 1024)])
def test_against_reference_build(desc, n0):
    o = Oracle(desc, n0)
    r = RefNFL(o.n, o.m * 60)
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.integers(0, p, o.n, dtype=np.uint64) for p in o.p])
    assert np.array_equal(o.ntt_powphi(x), r.ntt_powphi(x))
    assert np.array_equal(o.invntt(x), r.invntt(x))
    data = rng.integers(0, 256, 3 * 8192, dtype=np.uint8)
    assert np.array_equal(o.import_element(data), r.import_buffers(data, 3, o.A))
    nl = 33
    db = np.concatenate([o.import_element(rng.integers(0, 256, 2000, dtype=np.uint8)) for _ in range(nl)])
    P = o.polys(2000)
    q = np.concatenate([rng.integers(0, p, o.n, dtype=np.uint64) for _ in range(nl) for _ in range(2) for p in o.p])
    assert np.array_equal(o.reply_level(db, q, nl, 1, P), r.reply_level(db, q, nl, 1, P))


def test_planted_mistakes_are_caught():
    """Mutated restatements (in Python, on the CPU) must disagree with the oracle/golden."""
    g = GOLD["c1_plain"]
    o = Oracle(g["crypto"], g["dims"][0])
    rng = np.random.default_rng(2)
    p = o.p[0]
    nl = 200
    x = rng.integers(0, 2 * p, nl, dtype=np.uint64)
    y = rng.integers(0, p, nl, dtype=np.uint64)
    L = Oracle.lib()

    def chain(order, lazy=True):
        acc = 0
        for e in order:
            yq = L.xo_shoup_quot(int(y[e]), p)
            r = (int(x[e]) * int(y[e]) - ((int(x[e]) * yq) >> 64) * p) % 2**64
            acc = (acc + r) % 2**64
            if lazy and acc >= p:
                acc -= p
        return acc

    ref = 0
    for e in range(nl):
        ref = L.xo_mac_step(ref, int(x[e]), int(y[e]), L.xo_shoup_quot(int(y[e]), p), p)
    assert chain(range(nl)) == ref
    # reversed element order: same residues, different lazy words somewhere in a reply
    o2 = Oracle(C2, 64)
    rr = np.random.default_rng(4)
    n2 = 48
    db2 = np.concatenate([rr.integers(0, 2 * pp, o2.n, dtype=np.uint64) for _ in range(n2) for pp in o2.p])
    q2 = np.concatenate([rr.integers(0, pp, o2.n, dtype=np.uint64) for _ in range(n2) for _ in range(2) for pp in o2.p])
    fwd = o2.reply_level(db2, q2, n2, 1, 1)
    rev = o2.reply_level(db2.reshape(n2, -1)[::-1].copy(), q2.reshape(n2, -1)[::-1].copy(), n2, 1, 1)
    pv = np.tile(np.repeat(np.array(o2.p, dtype=np.uint64), o2.n), 2)
    assert np.array_equal(fwd % pv, rev % pv)
    assert np.any(fwd != rev)
    # a canonical (fully reduced) accumulator misses the golden's non-canonical words
    assert g["noncanonical_words"] > 0
    # transposed twiddle index: the NTT residue check fails
    n = o.n
    w = o.tables("w")
    xx = np.array(rng.integers(0, p, n, dtype=np.uint64))
    good = o.ntt_powphi(xx)
    bad_w = w.copy()
    bad_w[1], bad_w[2] = w[2], w[1]
    import ctypes as C
    ptr = o.tab.w
    arr = np.ctypeslib.as_array(ptr, shape=(o.mn,))
    saved = arr.copy()
    arr[:] = bad_w
    try:
        assert not np.array_equal(o.ntt_powphi(xx), good)
    finally:
        arr[:] = saved
    # a bit window shifted by one bit changes the absorbed chunks
    data = o.rand_bytes(5, 0, 4096)
    big = int.from_bytes(data.tobytes(), "little")
    out = o.absorb(data)
    assert int(out[1]) == (big >> o.A) & ((1 << o.A) - 1)
    assert int(out[1]) != (big >> (o.A + 1)) & ((1 << o.A) - 1) or int(out[2]) != (big >> (2 * o.A + 1)) & ((1 << o.A) - 1)


@pytest.mark.parametrize("vec", GOLD["alpha"], ids=lambda v: f"{v['crypto']}-{v['n_elems']}by{v['alpha']}")
def test_golden_alpha_aggregation(vec):
    """The oracle's alpha aggregation (and import) against the reference's
    readAggregatedStream + deserializeDataNFL bytes; zero-padding the last unit instead
    (the plausible 'fixed' reading) must not match."""
    o = Oracle(vec["crypto"], vec["n0"])
    n, L, a = vec["n_elems"], vec["elem_bytes"], vec["alpha"]
    data = o.rand_bytes(vec["seed"], 0, n * L)
    units = Oracle.aggregate_units(data, n, L, a)
    assert units.shape[0] == vec["units"]
    assert hashlib.sha256(units.tobytes()).hexdigest() == vec["units_sha256"]
    words = np.concatenate([o.import_element(u) for u in units])
    assert hashlib.sha256(words.tobytes()).hexdigest() == vec["ntt_sha256"]
    zp = Oracle.aggregate_units(data, n, L, a, zero_pad=True)
    assert hashlib.sha256(zp.tobytes()).hexdigest() != vec["units_sha256"]


@pytest.mark.skipif(not RefNFL.available(), reason="reference build (oracle/_ref) absent")
def test_alpha_aggregation_against_reference_build():
    """Every (n_elems, alpha, element size) up to 13 x 5: the oracle's units equal the
    reference's DBHandler/DBVectorProcessor output, ragged element sizes included."""
    rng = np.random.default_rng(12)
    for n in range(1, 14):
        for a in range(1, 6):
            sizes = rng.integers(1, 9, n)
            els = [rng.integers(0, 256, int(s), dtype=np.uint8).tobytes() for s in sizes]
            L = int(sizes.max())
            padded = np.concatenate([np.frombuffer(e + bytes(L - len(e)), dtype=np.uint8) for e in els])
            assert np.array_equal(Oracle.aggregate_units(padded, n, L, a), RefNFL.aggregate(els, a)), (n, a)


@pytest.mark.skipif(not RefNFL.available(), reason="reference build (oracle/_ref) absent")
@pytest.mark.parametrize("desc,dims,n,L", [(C3, [16, 12], 180, 1024), (C2, [20, 9], 171, 16384)])
def test_row_streamed_level0_matches_whole_database(desc, dims, n, L):
    """The row-streamed level 0 behind the full-size goldens (one group's elements made,
    imported and accumulated at a time) equals the reference's arithmetic over the whole
    imported database, padding units and partial group ranges included."""
    o = Oracle(desc, dims[0])
    r = RefNFL(o.n, o.m * 60)
    P0 = o.polys(L)
    data = o.rand_bytes(5, 0, n * L)
    words = np.zeros(dims[0] * dims[1] * P0 * o.mn, dtype=np.uint64)
    for e in range(n):
        words[e * P0 * o.mn:(e + 1) * P0 * o.mn] = r.import_buffers(data[e * L:(e + 1) * L], 1, o.A)
    q = o.make_query(o.keygen(3), dims, 7, 9)
    want = r.reply_level(words, q[0], dims[0], dims[1], P0, 4)
    assert np.array_equal(r.level0_stream(o.A, 5, n, L, dims[0], 0, dims[1], P0, q[0], 4), want)
    parts = [r.level0_stream(o.A, 5, n, L, dims[0], g, min(g + 5, dims[1]), P0, q[0], 2)
             for g in range(0, dims[1], 5)]
    assert np.array_equal(np.concatenate(parts), want)
    assert np.array_equal(o.reply_level(words, q[0], dims[0], dims[1], P0), want)


def test_full_size_golden_entries():
    """The full-size fixtures describe the configurations they claim (A, polys, grids)."""
    from xpir_amd.workloads import C3_RECURSIVE, batch_databases, serving_d1_views
    c3 = GOLD["c3_recursive"]
    assert (c3["A"], c3["P0"], c3["Pd"], c3["dims"], c3["n_elems"]) == (74, 1, 6, [1024, 1024], 1 << 20)
    assert c3["index"] == C3_RECURSIVE["index"]
    for g, spec in zip(GOLD["c4_batched"], batch_databases()):
        o = Oracle(spec["crypto"], spec["dims"][0])
        assert (g["dims"], g["n_elems"], g["elem_bytes"], g["A"]) == (spec["dims"], spec["n_elems"],
                                                                      spec["elem_bytes"], o.A)
        assert g["P0"] == o.polys(spec["elem_bytes"]) and g["Pd"] == o.next_polys(g["P0"])
    for g, v in zip(GOLD["c5_d1_views"], serving_d1_views()):
        assert (g["dims"], g["index"], g["noise_seed"]) == (v["dims"], v["index"], v["noise_seed"])
