// ntt_dif.hpp -- the reference's negacyclic DIF NTT core on n words in LDS
// (crypto/NFLlib.hpp:450-537): Harvey's lazy radix-2 butterflies on [0, 2p) words, then
// the fused final radix-4 block.  Shared by the import kernel (forward transform) and the
// decryption kernel (the same DIF with the inverse root, between bit-reversal permutations).
#pragma once
#include "arith.hpp"

// Radix-2 DIF layers L = 0 .. log2(n) - 3 over sh[0..n), for a workgroup of any size; the
// twiddles of the layer with half-block h start at w[n - 2h].  Ends with a barrier.
__device__ __forceinline__ void dif_radix2_layers(uint64_t *sh, uint32_t n, uint32_t logn, const uint64_t *w,
                                                  const uint64_t *wq, uint64_t p, uint32_t delta) {
  const uint64_t p2 = 2 * p;
  // radix-2 DIF layers (NFLlib.hpp:478-499), two per pass: a thread takes the four words
  // {i, i + h/2, i + h, i + 3h/2} of a 2h-block, which the butterflies of layer L
  // (pairs at distance h) and of layer L + 1 (distance h/2) exchange only among
  // themselves.  Every butterfly is the reference's, on the same inputs and twiddles.
  const uint32_t radix2_layers = logn - 2;
  uint32_t L = 0, lg_half = logn - 1;
  for (; L + 1 < radix2_layers; L += 2, lg_half -= 2) {
    const uint32_t half = 1u << lg_half, q = half >> 1;
    const uint64_t *tw = w + (n - 2 * half), *twq = wq + (n - 2 * half);  // layer L
    const uint64_t *tv = w + (n - half), *tvq = wq + (n - half);          // layer L + 1
    for (uint32_t b = threadIdx.x; b < n / 4; b += blockDim.x) {
      const uint32_t i = b & (q - 1);
      const uint32_t e0 = ((b >> (lg_half - 1)) << (lg_half + 1)) + i;
      uint64_t x0 = sh[e0], x1 = sh[e0 + q], x2 = sh[e0 + half], x3 = sh[e0 + half + q];
      // layer L: (x0, x2) with twiddle i, (x1, x3) with twiddle i + q
      uint64_t s = x0 + x2;
      s = s >= p2 ? s - p2 : s;
      x2 = shoup_lazy_sf(x0 - x2 + p2, tw[i], twq[i], delta);
      x0 = s;
      s = x1 + x3;
      s = s >= p2 ? s - p2 : s;
      x3 = shoup_lazy_sf(x1 - x3 + p2, tw[i + q], twq[i + q], delta);
      x1 = s;
      // layer L + 1: (x0, x1) and (x2, x3), both with twiddle i
      s = x0 + x1;
      s = s >= p2 ? s - p2 : s;
      x1 = shoup_lazy_sf(x0 - x1 + p2, tv[i], tvq[i], delta);
      x0 = s;
      s = x2 + x3;
      s = s >= p2 ? s - p2 : s;
      x3 = shoup_lazy_sf(x2 - x3 + p2, tv[i], tvq[i], delta);
      x2 = s;
      sh[e0] = x0;
      sh[e0 + q] = x1;
      sh[e0 + half] = x2;
      sh[e0 + half + q] = x3;
    }
    __syncthreads();
  }
  if (L < radix2_layers) {  // an odd layer count leaves one single layer
    const uint32_t half = 1u << lg_half;
    const uint64_t *tw = w + (n - 2 * half), *twq = wq + (n - 2 * half);
    for (uint32_t b = threadIdx.x; b < n / 2; b += blockDim.x) {
      const uint32_t j = b & (half - 1);
      const uint32_t i0 = ((b >> lg_half) << (lg_half + 1)) + j, i1 = i0 + half;
      const uint64_t u0 = sh[i0], u1 = sh[i1];
      uint64_t s = u0 + u1;
      s = s >= p2 ? s - p2 : s;
      const uint64_t d = u0 - u1 + p2;
      sh[i0] = s;
      sh[i1] = shoup_lazy_sf(d, tw[j], twq[j], delta);
    }
    __syncthreads();
  }

}

// The last two layers on four consecutive words (u0..u3), the reference's fused block:
// W is the layer-(log2 n - 2) twiddle w[n - 4 + 1].
__device__ __forceinline__ void dif_last_two(uint64_t (&z)[4], uint64_t u0, uint64_t u1, uint64_t u2, uint64_t u3,
                                             uint64_t W, uint64_t Wq, uint64_t p, uint32_t delta) {
  const uint64_t p2 = 2 * p;
  uint64_t v0 = u0 + u2; v0 = v0 >= p2 ? v0 - p2 : v0;
  uint64_t v2 = u0 - u2; v2 = (int64_t)v2 < 0 ? v2 + p2 : v2;
  uint64_t v1 = u1 + u3; v1 = v1 >= p2 ? v1 - p2 : v1;
  const uint64_t v3 = shoup_lazy_sf(u1 - u3 + p2, W, Wq, delta);
  z[0] = v0 + v1; z[0] = z[0] >= p2 ? z[0] - p2 : z[0];
  z[1] = v0 - v1; z[1] = (int64_t)z[1] < 0 ? z[1] + p2 : z[1];
  z[2] = v2 + v3; z[2] = z[2] >= p2 ? z[2] - p2 : z[2];
  z[3] = v2 - v3; z[3] = (int64_t)z[3] < 0 ? z[3] + p2 : z[3];
}
