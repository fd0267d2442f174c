"""Workload definitions (seeds and sizes), shared by tests and bench.py.

Conventions fixed here (the task's configurations leave them open):
  * "10 KB" / "4 KB" / "1 KB" element sizes are KiB (10240 / 4096 / 1024 bytes);
    element counts such as "10k" are decimal.
  * DB bytes: byte i of the database is byte (i mod 8) of word i//8 of the seeded
    SplitMix64 stream `db_seed` (xpir_amd/csrc/arith.hpp: rand_word).
  * The query index is word 0 of stream `index_seed`, modulo the element count.
  * Secret key: stream `key_seed`; noise and uniform `a`: stream `noise_seed`
    (and noise_seed ^ 0xA0A0A0A0A0A0A0A0 for `a`).
"""


def _index(seed, n):
    z = (seed + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    z ^= z >> 31
    return z % n


C1_PLAIN = dict(name="plain-LWE:80:1024:60-1000x10KiB", crypto="LWE:80:1024:60", n_elems=1000,
                elem_bytes=10240, dims=[1000], db_seed=1, index=_index(2, 1000), noise_seed=3, key_seed=4)

C2_SINGLE = dict(name="single-LWE:80:2048:120-65536x4KiB", crypto="LWE:80:2048:120", n_elems=65536,
                 elem_bytes=4096, dims=[65536], db_seed=21, index=_index(22, 65536), noise_seed=23,
                 key_seed=24)

C3_RECURSIVE = dict(name="recursive-LWE:100:4096:180-2^20x1KiB-d2", crypto="LWE:100:4096:180",
                    n_elems=1 << 20, elem_bytes=1024, dims=[1024, 1024], db_seed=31,
                    index=_index(32, 1 << 20), noise_seed=33, key_seed=34)

# scaled-down grid of the recursive configuration: same cryptosystem and element size
C3_SCALED = dict(C3_RECURSIVE, name="recursive-scaled-64x64", n_elems=4096, dims=[64, 64],
                 index=_index(32, 4096))


def dim_sizes(n: int, alpha: int, d: int):
    """The d-dimensional grid for n elements: factors floor((n/alpha)^(1/d)), then each
    incremented in turn until their product covers n/alpha
    (PIROptimizer::getDimSize, apps/optim/PIROptimizer.cpp:338-366)."""
    import math
    new_n = math.ceil(n / alpha)
    f = [math.floor(new_n ** (1.0 / d))] * d
    if float(f[0]) != new_n ** (1.0 / d):
        prod, j = 1, 0
        while prod < new_n and j < d:
            prod = 1
            f[j] += 1
            j += 1
            for x in f:
                prod *= x
    return f


# Batched uneven: 8 resident LWE:80:2048:120 databases; element sizes cycle through
# {256 B, 1 KiB, 4 KiB, 16 KiB}; d = 2 grids (d = 1 queries on these sizes would be
# 0.6-61 GiB each); 1..64 queries per database, counts from seed 41.
BATCH_COUNTS = [10_000, 25_000, 50_000, 100_000, 200_000, 400_000, 700_000, 1_000_000]
BATCH_SIZES = [256, 1024, 4096, 16384]


def batch_databases(scale: float = 1.0):
    out = []
    for i, n in enumerate(BATCH_COUNTS):
        n = max(4, int(n * scale))
        out.append(dict(crypto="LWE:80:2048:120", n_elems=n, elem_bytes=BATCH_SIZES[i % 4],
                        dims=dim_sizes(n, 1, 2), db_seed=4100 + i, key_seed=4200 + i))
    return out


def batch_golden_query(i: int, spec: dict) -> dict:
    """The seeded query whose reply the golden fixtures hold for batched database i."""
    return dict(index=_index(4300 + i, spec["n_elems"]), noise_seed=4400 + i, key_seed=spec["key_seed"])


def serving_d1_views():
    """d = 1 views of the two smallest batched databases (their d = 1 queries, 625 MiB and
    1.5 GiB, fit a serving budget; A depends on the first dimension, so each view is a
    transform of its own), with the seeded query the golden fixtures answer."""
    out = []
    for i, spec in enumerate(batch_databases()[:2]):
        v = dict(spec, dims=[spec["n_elems"]], name=f"serving-d1-db{i}-{spec['n_elems']}x{spec['elem_bytes']}")
        v.update(index=_index(4500 + i, spec["n_elems"]), noise_seed=4600 + i)
        out.append(v)
    return out


def batch_query_counts(seed: int = 41, n_db: int = 8, lo: int = 1, hi: int = 64):
    import numpy as np
    return [int(x) for x in np.random.default_rng(seed).integers(lo, hi + 1, size=n_db)]
