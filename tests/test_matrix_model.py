"""The matrix-core level 0 (xpir_amd/csrc/mx.hip) on the CPU: its exactness argument
against the reference's chain step, and its digit and lane arithmetic.

1. A chain (NFLlib.hpp:303-321, here the oracle's xo_mac_step) over n elements ends at the
   word the matrix path builds: the residue of the first n - 40 products' sum, then the
   last 40 steps with the reference's step from that residue, whenever the tail keeps at
   least 8 steps without subtracting p.  The claim behind it (max-plus algebra of the
   excess multiple j of p): every start j0 <= kept ends where the start at 0 does, checked
   here in exact integers for every j0 <= 8; a start j0 > kept is shown to differ.
2. Signed byte digits of a word reconstruct it; the epilogue's lane-by-lane reduction of
   the 32 x 32 int32 products (mx.hip: tile_epilogue) equals the weighted digit sum mod p.
"""
import random

import pytest

from oracle.oracle import Oracle

M64 = (1 << 64) - 1
TAIL = 40


def _lazy(x, y, yq, p):
    return (x * y - ((x * yq) >> 64) * p) & M64


def _tail(start, xs, ys, p, exact=False):
    """The reference's step over the tail from acc = start; the final word and the count of
    steps that kept acc + r (did not subtract p).  exact: without the 64-bit wrap (the
    model's integers)."""
    acc, kept = start, 0
    for x, y in zip(xs, ys):
        s = acc + _lazy(x, y, ((y << 64) // p) & M64, p)
        s = s if exact else s & M64
        if s >= p:
            acc = s - p
        else:
            acc, kept = s, kept + 1
    return acc, kept


def _chain(L, xs, ys, p):
    acc = 0
    for x, y in zip(xs, ys):
        acc = L.xo_mac_step(acc, x, y, ((y << 64) // p) & M64, p)
    return acc


@pytest.mark.parametrize("kind", ["random", "lazy_edges", "wrapping"])
def test_prefix_residue_and_tail_give_the_chain_word(kind):
    L = Oracle.lib()
    rng = random.Random(hash(kind) & 0xFFFF)
    settled = unsettled = 0
    for trial in range(300):
        p = L.xo_modulus(trial % 17)
        n = rng.randrange(TAIL + 32, 260)
        if kind == "random":
            xs = [rng.randrange(2 * p) for _ in range(n)]
            ys = [rng.randrange(p) for _ in range(n)]
        elif kind == "lazy_edges":
            xs = [rng.choice([0, 1, p - 1, p, 2 * p - 1, rng.randrange(2 * p)]) for _ in range(n)]
            ys = [rng.choice([0, 1, p - 1, rng.randrange(p)]) for _ in range(n)]
        else:  # every product p - 1: the residue wraps at almost every step
            xs, ys = [1] * n, [p - 1] * n
        want = _chain(L, xs, ys, p)
        e0 = n - TAIL
        rho = sum(x * y for x, y in zip(xs[:e0], ys[:e0])) % p
        got, kept = _tail(rho, xs[e0:], ys[e0:], p)
        if kept >= 8:
            settled += 1
            assert got == want, (kind, trial)
            # every start with the same residue and at most kept excess multiples of p ends
            # at the same word (exact integers: the model has no wrap)
            for j0 in range(1, 9):
                assert _tail(rho + j0 * p, xs[e0:], ys[e0:], p, exact=True)[0] == got
        else:
            unsettled += 1  # the kernel recomputes such a chain whole
    if kind == "wrapping":
        assert settled == 0 and unsettled == 300
    else:
        assert settled >= 290


def test_start_beyond_kept_steps_differs():
    """A start with more excess multiples of p than the tail's kept steps can end elsewhere:
    the bound kept >= j0 is tight, so the tail length matters."""
    L = Oracle.lib()
    rng = random.Random(5)
    found = False
    for trial in range(3000):
        p = L.xo_modulus(trial % 17)
        n = rng.randrange(1, 8)
        xs = [rng.randrange(2 * p) for _ in range(n)]
        ys = [rng.randrange(p) for _ in range(n)]
        rho = rng.randrange(p)
        got, kept = _tail(rho, xs, ys, p, exact=True)
        if _tail(rho + (kept + 1) * p, xs, ys, p, exact=True)[0] != got:
            found = True
            break
    assert found


def _digits(x):
    b = ((x + 0x8080808080808080) & M64) ^ 0x8080808080808080
    return [((b >> (8 * j)) & 0xFF) - 256 * (((b >> (8 * j)) & 0xFF) >= 128) for j in range(8)]


def test_signed_digits_reconstruct_words():
    rng = random.Random(1)
    for _ in range(5000):
        x = rng.choice([rng.getrandbits(62), 0, (1 << 62) - 1, 0x7F7F7F7F7F7F7F7F])
        d = _digits(x)
        assert all(-128 <= v <= 127 for v in d) and sum(v << (8 * j) for j, v in enumerate(d)) == x


def test_epilogue_lane_reduction():
    """The tile epilogue's steps on the instruction's output layout (lane L: column L % 32,
    rows 8(v/4) + 4(L/32) + v%4 in register v): per row group g and unit ul, the sum of
    2^(8(i+j)) S[8g+i][8ul+j] mod p."""
    L = Oracle.lib()
    rng = random.Random(9)
    for t in range(12):
        p = L.xo_modulus(t)
        delta = (1 << 61) - p
        S = [[rng.randint(-2 ** 29, 2 ** 29) for _ in range(32)] for _ in range(32)]
        acc = {ln: [S[8 * (v // 4) + 4 * (ln >> 5) + v % 4][ln & 31] for v in range(16)] for ln in range(64)}
        D = {ln: [sum(acc[ln][4 * g + b] << (8 * b) for b in range(4)) for g in range(4)] for ln in range(64)}
        w0, w1 = {}, {}
        for ln in range(64):  # xor 1, exact: P = D(even lane) + 2^8 D(odd lane)
            odd, pr = ln & 1, ln ^ 1
            e, o = (pr, ln) if odd else (ln, pr)
            pw = pow(2, 8 * (4 * (ln >> 5) + (ln & 6)), p)
            pq = (pw << 64) // p
            for k, g in enumerate((odd, 2 + odd)):
                P = D[e][g] + (D[o][g] << 8)
                assert abs(P) < 2 ** 61.02
                (w0 if k == 0 else w1)[ln] = _lazy((P + 4 * p) & M64, pw, pq, p)  # in [0, 2p)
        v = {}
        for ln in range(64):  # xor 2, 4 (below 8p), fold, xor 32
            hi2 = (ln >> 1) & 1
            kept = w1 if hi2 else w0  # group odd (bit 1 clear) or 2 + odd; the partner sends the same group
            v[ln] = kept[ln] + kept[ln ^ 2]
        v = {ln: v[ln] + v[ln ^ 4] for ln in range(64)}
        assert all(x < 1 << 64 for x in v.values())
        v = {ln: ((x & ((1 << 61) - 1)) + (x >> 61) * delta) % p for ln, x in v.items()}
        v = {ln: (v[ln] + v[ln ^ 32]) % p for ln in range(64)}
        for ln in range(32):
            if (ln & 7) < 4:
                g, ul = ln & 7, (ln >> 3) & 3
                want = sum(S[8 * g + i][8 * ul + j] << (8 * (i + j)) for i in range(8) for j in range(8)) % p
                assert v[ln] == want
