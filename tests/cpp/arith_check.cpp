// arith_check.cpp -- host-side check of the special-form arithmetic the MAC and NTT
// kernels use (xpir_amd/csrc/arith.hpp, __host__ __device__ functions), against the
// textbook forms with 128-bit integers:
//   shoup_quot(y)            == floor(y * 2^64 / p)                    (mod 2^64)
//   shoup_lazy_sf(x, y, y')  == x*y - floor(x*y' / 2^64)*p            (mod 2^64)
// for every modulus given on the command line, on boundary words and on random words.
// usage: arith_check COUNT p1 [p2 ...]    (prints "ok" and exits 0 when every check holds)
#include <stdio.h>
#include <stdlib.h>

#include "../../xpir_amd/csrc/arith.hpp"

typedef unsigned __int128 u128;

static uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  const uint64_t count = strtoull(argv[1], 0, 10);
  uint64_t failures = 0, checks = 0;
  for (int a = 2; a < argc; a++) {
    const uint64_t p = strtoull(argv[a], 0, 10);
    const uint32_t delta = (uint32_t)((1ULL << 61) - p);
    const uint64_t edges[] = {0, 1, 2, p - 1, p, p + 1, 2 * p - 1, 2 * p, 4 * p - 1, (1ULL << 61) - 1, 1ULL << 61,
                              (1ULL << 62) - 1, 1ULL << 62, (1ULL << 63) - 1, 1ULL << 63, ~0ULL, ~0ULL - 1};
    const int ne = sizeof(edges) / sizeof(edges[0]);
    uint64_t s = p;
    for (uint64_t k = 0; k < count + (uint64_t)(ne * ne); k++) {
      uint64_t x, y;
      if (k < (uint64_t)(ne * ne)) {
        x = edges[k / ne];
        y = edges[k % ne];
      } else {
        x = splitmix(s);
        y = splitmix(s);
        switch (k % 4) {  // the ranges the kernels see: lazy DB words, canonical query words
          case 0: x %= 2 * p; y %= p; break;
          case 1: x >>= 2; break;
          case 2: y %= p; break;
          default: break;
        }
      }
      const uint64_t yq = (uint64_t)(((u128)y << 64) / p);
      checks++;
      if (shoup_quot(y, p, delta) != yq) {
        if (failures++ < 5) fprintf(stderr, "shoup_quot p=%llu y=%llu\n", (unsigned long long)p, (unsigned long long)y);
      }
      const uint64_t want = x * y - (uint64_t)(((u128)x * yq) >> 64) * p;
      if (shoup_lazy_sf(x, y, yq, delta) != want) {
        if (failures++ < 5)
          fprintf(stderr, "shoup_lazy_sf p=%llu x=%llu y=%llu\n", (unsigned long long)p, (unsigned long long)x,
                  (unsigned long long)y);
      }
      if (shoup_lazy_qf(x, y, p, delta) != want) failures++;
    }
  }
  if (failures) {
    fprintf(stderr, "%llu failures in %llu checks\n", (unsigned long long)failures, (unsigned long long)checks);
    return 1;
  }
  printf("ok %llu\n", (unsigned long long)checks);
  return 0;
}
