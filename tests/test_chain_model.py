"""CPU model of the two chain decompositions the GPU MAC kernels use, checked against the
sequential lazy chain of NFLlib::mulandaddShoup (crypto/NFLlib.hpp:303-321): starting
from zero, `acc += r; if acc >= p: acc -= p` for lazy products r in [0, 2p).

* segmented kernel: acc = rho + j*p with rho canonical; each step is the max-plus map
  j -> max(j + d, 0), d = floor((rho + r) / p) - 1; maps compose as
  (D1, M1) then (D2, M2) = (D1 + D2, max(M1 + D2, M2)).
* streaming kernel: a block of steps is summarised without knowing its incoming residue
  (rho0 = A_S mod p, J_t = t - floor(A_t / p), Jend, Jmax, Rstar = min residue where J
  reaches Jmax); with the incoming residue the block's map is
  D = [rho_in + rho0 >= p] - Jend,  M = D + Jmax - [rho_in + Rstar >= p].

Pure-Python big integers, so no step can overflow; the chains mix uniform lazy products
with the boundary values p - 1, p, p + 1, 2p - 1 and runs of products at or above p.
"""
import random

import pytest

MODULI = [2305843009213317121, 2305843009213120513, 2305843009206845441]


def sequential(rs, p):
    acc = 0
    for r in rs:
        acc += r
        if acc >= p:
            acc -= p
    return acc


def maxplus_chain(rs, p, cuts):
    """Segmented form: per-tile maps from the known incoming residue, folded in order."""
    rho, j = 0, 0
    prev = 0
    for c in cuts + [len(rs)]:
        D, M = 0, None
        for r in rs[prev:c]:
            k = (rho + r) // p
            rho = (rho + r) - k * p
            d = k - 1
            D += d
            M = max((M if M is not None else -10**9) + d, 0)
        if M is not None:
            j = max(j + D, M)
        prev = c
    return rho + j * p


def block_summary(rs, p):
    rho, J, Jmax, Rs = 0, 0, None, None
    for r in rs:
        u = rho + r
        k = u // p
        rho = u - k * p
        J += 1 - k
        if Jmax is None or J > Jmax or (J == Jmax and rho < Rs):
            Jmax, Rs = J, rho
    return rho, J, Jmax, Rs


ZBIAS = 1 << 21  # blocks hold at most 2^20 elements: |Z| < 2^81 < ZBIAS * p


def z_to_j_r(z, p):
    """J = ceil(z / p) and R = J * p - z in [0, p), for |z| < ZBIAS * p, as the kernel
    computes them: N = -z + ZBIAS * p >= 0 split as H * 2^61 + L = H * p + (H * delta + L)."""
    delta = (1 << 61) - p
    N = -z + ZBIAS * p
    assert 0 <= N < 1 << 96
    H, L = N >> 61, N & ((1 << 61) - 1)
    S = H * delta + L
    b = 1 if S >= p else 0
    q = H + b - ZBIAS  # floor(-z / p)
    return -q, S - b * p


def block_summary_z(rs, p):
    """The streaming kernel's per-step state: Z_t = t * p - A_t (= J_t * p - R_t), one add
    of p - r per step, and its running maximum; (J, R) pairs recovered at the block end.
    max_t Z_t is reached exactly where J_t = Jmax with the smallest residue, so it gives
    (Jmax, Rstar)."""
    z, zmax = 0, None
    for r in rs:
        z += p - r
        assert -(1 << 95) <= z < 1 << 95  # the kernel's 96-bit word
        if zmax is None or z > zmax:
            zmax = z
    Jend, rho0 = z_to_j_r(z, p)
    Jmax, Rs = z_to_j_r(zmax, p)
    return rho0, Jend, Jmax, Rs


def streaming_chain(rs, p, cuts):
    """Streaming form: block summaries computed blind, then scanned with the residues."""
    rho_in, j = 0, 0
    prev = 0
    for c in cuts + [len(rs)]:
        rho0, Jend, Jmax, Rs = block_summary(rs[prev:c], p)
        D = (1 if rho_in + rho0 >= p else 0) - Jend
        M = D + Jmax - (1 if rho_in + Rs >= p else 0)
        j = max(j + D, M)
        rho_in = (rho_in + rho0) % p
        prev = c
    return rho_in + j * p


def special_form_step(rho, r, p):
    """The kernels' step arithmetic: rho + r = h * 2^61 + l = h * p + v, v = l + h * delta."""
    delta = (1 << 61) - p
    u = rho + r
    h = u >> 61
    v = (u & ((1 << 61) - 1)) + h * delta
    b = 1 if v >= p else 0
    return v - b * p, h + b


def _chains(seed, count):
    rng = random.Random(seed)
    for t in range(count):
        p = MODULI[t % len(MODULI)]
        n = rng.randint(1, 200)
        kind = t % 4
        rs = []
        for _ in range(n):
            if kind == 0:
                r = rng.randrange(p) if rng.random() < 0.97 else rng.randrange(p, 2 * p)
            elif kind == 1:
                r = rng.choice([0, 1, p - 1, p, p + 1, 2 * p - 1, rng.randrange(2 * p)])
            elif kind == 2:
                r = rng.randrange(p, 2 * p) if rng.random() < 0.5 else rng.randrange(p)
            else:
                r = rng.randrange(2 * p)
            rs.append(r)
        k = rng.randint(0, min(n - 1, 10))
        cuts = sorted(rng.sample(range(1, n), k)) if n > 1 else []
        yield p, rs, cuts


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_segmented_and_streaming_forms_equal_sequential_chain(seed):
    for p, rs, cuts in _chains(seed, 1500):
        want = sequential(rs, p)
        assert maxplus_chain(rs, p, cuts) == want
        assert streaming_chain(rs, p, cuts) == want


@pytest.mark.parametrize("seed", [5, 6])
def test_running_max_form_equals_block_summary(seed):
    for p, rs, cuts in _chains(seed, 1500):
        assert block_summary_z(rs, p) == block_summary(rs, p)
    rng = random.Random(seed)
    for p in MODULI:  # long blocks at the extremes of the 96-bit range
        for rs in ([0] * 4096, [2 * p - 1] * 4096, [p] * 3000, [rng.choice([0, 2 * p - 1]) for _ in range(5000)]):
            assert block_summary_z(rs, p) == block_summary(rs, p)


def test_z_to_j_r_at_the_edges():
    for p in MODULI:
        for z in (0, 1, -1, p - 1, p, p + 1, -p, -p + 1, -p - 1, (1 << 81) - 1, -(1 << 81), 5 * p, -7 * p + 3):
            J, R = z_to_j_r(z, p)
            assert 0 <= R < p and J * p - z == R


def test_special_form_step_is_division_by_p():
    rng = random.Random(9)
    for t in range(20000):
        p = MODULI[t % len(MODULI)]
        rho = rng.choice([0, 1, p - 1, rng.randrange(p)])
        r = rng.choice([0, 1, p - 1, p, p + 1, 2 * p - 1, rng.randrange(2 * p)])
        rho2, k = special_form_step(rho, r, p)
        assert (rho2, k) == ((rho + r) % p, (rho + r) // p)


def test_planted_mistake_in_streaming_form_is_caught():
    """Dropping the Rstar correction (taking M = D + Jmax) must break some chain."""
    def wrong(rs, p, cuts):
        rho_in, j, prev = 0, 0, 0
        for c in cuts + [len(rs)]:
            rho0, Jend, Jmax, Rs = block_summary(rs[prev:c], p)
            D = (1 if rho_in + rho0 >= p else 0) - Jend
            j = max(j + D, D + Jmax)
            rho_in = (rho_in + rho0) % p
            prev = c
        return rho_in + j * p
    assert any(wrong(rs, p, cuts) != sequential(rs, p) for p, rs, cuts in _chains(4, 600))
