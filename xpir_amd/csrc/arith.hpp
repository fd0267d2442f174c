// arith.hpp -- exact 64-bit modular arithmetic for the NFLLWE reply path on CDNA4.
//
// All words are uint64; moduli are the reference's p_i = 2^61 - delta_i with
// delta_i = c_i*2^14 - 1 < 2^21 (crypto/NFLParams.cpp:34).  The lazy conventions
// follow crypto/NFLlib.hpp:255-321 exactly, because the reply bytes depend on them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define XPG_MAX_M 4

struct ModSet {
  uint64_t p[XPG_MAX_M];
  uint32_t delta[XPG_MAX_M];  // 2^61 - p (below 2^21)
};

__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// x*y - floor(x*yq / 2^64)*p modulo 2^64; in [0, 2p) for p < 2^62 and any x < 2^64
// (crypto/NFLlib.hpp:283-285 before its final conditional subtraction).
__device__ __forceinline__ uint64_t shoup_lazy(uint64_t x, uint64_t y, uint64_t yq, uint64_t p) {
  return x * y - mulhi64(x, yq) * p;
}

// mulmodShoup: the lazy product with one conditional subtraction (canonical for x < 2^64).
__device__ __forceinline__ uint64_t shoup_canon(uint64_t x, uint64_t y, uint64_t yq, uint64_t p) {
  uint64_t r = shoup_lazy(x, y, yq, p);
  return r >= p ? r - p : r;
}

// floor(y * 2^64 / p) mod 2^64 -- the Shoup quotient the reference precomputes per query
// word on pushQuery (PIRReplyGeneratorNFL_internal.cpp:760-767) -- derived without a
// division from p = 2^61 - delta:  y * 2^64 / p = 8y + 8*y*delta/p, and with
// t0 = (y*delta) >> 58 and N = 8*y*delta = t0 * 2^61 + nm = t0 * p + (nm + t0*delta),
// where nm + t0*delta < 2p:  y' = 8y + t0 + [nm + t0*delta >= p]  (mod 2^64).
// Holds for every y < 2^64 (checked exhaustively at the edges and on random words).
__host__ __device__ __forceinline__ uint64_t shoup_quot(uint64_t y, uint64_t p, uint32_t delta) {
  const uint32_t d = delta, yl = (uint32_t)y, yh = (uint32_t)(y >> 32);
  const uint64_t a0 = (uint64_t)yl * d;
  const uint64_t a1 = (uint64_t)yh * d + (a0 >> 32);  // y*delta = a1 * 2^32 + low word of a0
  const uint32_t t0 = (uint32_t)(a1 >> 26);
  const uint64_t nm = (((a1 & 0x3ffffffULL) << 32) | (uint32_t)a0) << 3;
  const uint32_t t = t0 + ((uint64_t)t0 * d + nm >= p ? 1u : 0u);  // floor(8*y*delta / p)
  return (y << 3) + t;
}

// The lazy Shoup product x*y - floor(x*yq/2^64)*p (mod 2^64) for p = 2^61 - delta,
// written as 32x32->64 multiply-adds (v_mad_u64_u32) and 32-bit low multiplies:
//   q = high word of x*yq,   r = x*y + q*delta - (q << 61)  (mod 2^64),
// where only the low three bits of q reach the high word through q << 61.
// Equal to shoup_lazy(x, y, yq, p) for every x, y, yq < 2^64.
__host__ __device__ __forceinline__ uint64_t shoup_lazy_sf(uint64_t x, uint64_t y, uint64_t yq, uint32_t delta) {
  const uint32_t d = delta, yl = (uint32_t)y, yh = (uint32_t)(y >> 32);
  const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32), ql = (uint32_t)yq, qh = (uint32_t)(yq >> 32);
  const uint64_t m1 = (uint64_t)xh * ql + (((uint64_t)xl * ql) >> 32);
  const uint64_t m2 = (uint64_t)xl * qh + (uint32_t)m1;
  const uint64_t q = (uint64_t)xh * qh + (m1 >> 32) + (m2 >> 32);
  const uint64_t lo = (uint64_t)xl * yl + (uint64_t)(uint32_t)q * d;
  const uint32_t hi = (uint32_t)(lo >> 32) + xl * yh + xh * yl + (uint32_t)(q >> 32) * d - ((uint32_t)q << 29);
  return ((uint64_t)hi << 32) | (uint32_t)lo;
}

// The lazy Shoup product with the query quotient derived on the fly (shoup_quot):
// equal to shoup_lazy(x, y, shoup_quot(y, p, delta), p) for every x, y < 2^64.
__host__ __device__ __forceinline__ uint64_t shoup_lazy_qf(uint64_t x, uint64_t y, uint64_t p, uint32_t delta) {
  return shoup_lazy_sf(x, y, shoup_quot(y, p, delta), delta);
}

// One reply-chain step (crypto/NFLlib.hpp:303-321): acc += lazy product; -p if >= p.
__device__ __forceinline__ uint64_t mac_step(uint64_t acc, uint64_t x, uint64_t y, uint64_t yq,
                                             uint64_t p, uint32_t delta) {
  acc += shoup_lazy_sf(x, y, yq, delta);
  return acc >= p ? acc - p : acc;
}

// Per-modulus constants of the asm chain step below (wave-uniform except plo/phi, which
// the select needs in VGPRs: a VOP3 instruction reads one SGPR at most on this GPU).
struct StepConsts {
  uint64_t negp;   // 2^64 - p
  uint32_t delta;  // 2^61 - p
  uint32_t k29;    // 0xE0000000 = high word of 2^64 - p, so lo(q)*k29 = -(lo(q) << 29) mod 2^32
  uint32_t plo, phi;
};

__device__ __forceinline__ StepConsts step_consts(uint64_t p, uint32_t delta) {
  StepConsts k;
  k.negp = 0 - p;
  k.delta = delta;
  k.k29 = 0xE0000000u;
  k.plo = (uint32_t)p;
  k.phi = (uint32_t)(p >> 32);
  asm volatile("" : "+v"(k.plo), "+v"(k.phi));  // the select's operands live in VGPRs
  return k;
}

// One reply-chain step (crypto/NFLlib.hpp:303-321) as straight-line VALU code:
//   acc = acc + (x*y - q*p),  q = hi64(x*yq);   if (acc >= p) acc -= p   (all mod 2^64)
// for every x, y, yq, acc < 2^64 and p = 2^61 - delta.  18 VALU instructions:
//   q:   t = hi32(xl*ql) (from C, as {t, 0}), A = xh*ql + t (below 2^64), B = xl*qh + A with
//        its carry c, q = xh*qh + {hi32(B), c} -- the middle product's carry comes from
//        v_mad_u64_u32's carry-out instead of being recombined from 32-bit halves;
//   r-p: x*y + q*(2^64 - p) - p (mod 2^64), with 2^64 - p = 0xE0000000 * 2^32 + delta:
//        lo = xl*yl - p + lo(q)*delta (two 64-bit multiply-adds), and the four products
//        that reach only the high word (xl*yh, xh*yl, hi(q)*delta, lo(q)*0xE0000000) are
//        summed by a multiply and three multiply-adds whose low words alone are kept;
//   chain: u = acc + r - p; the reference's acc + r >= p exactly when u < 2^64 - p, then
//        u is the new acc, else u + p.
// Scratch v[2:9] and vcc; an s_nop separates a VALU write of vcc from its read as a lane
// mask (a hazard on this GPU the compiler cannot see inside inline asm).
#define XPG_MAC_STEP_ASM                                \
  "v_mad_u64_u32 v[4:5], vcc, %[xh], %[ql], %[T]\n\t"   \
  "v_mad_u64_u32 v[6:7], vcc, %[xl], %[qh], v[4:5]\n\t" \
  "v_mul_lo_u32 v2, %[xl], %[yh]\n\t"                   \
  "v_mov_b32 v8, v7\n\t"                                \
  "v_cndmask_b32_e64 v9, 0, 1, vcc\n\t"                 \
  "v_mad_u64_u32 v[4:5], vcc, %[xh], %[qh], v[8:9]\n\t" \
  "v_mad_u64_u32 v[2:3], vcc, %[xh], %[yl], v[2:3]\n\t" \
  "v_mad_u64_u32 v[6:7], vcc, %[xl], %[yl], %[negp]\n\t" \
  "v_mad_u64_u32 v[6:7], vcc, v4, %[d], v[6:7]\n\t"     \
  "v_mad_u64_u32 v[2:3], vcc, v5, %[d], v[2:3]\n\t"     \
  "v_mad_u64_u32 v[2:3], vcc, v4, %[k29], v[2:3]\n\t"   \
  "v_add_u32 v7, v7, v2\n\t"                            \
  "v_lshl_add_u64 %[acc], %[acc], 0, v[6:7]\n\t"        \
  "v_cmp_gt_u64_e64 vcc, %[negp], %[acc]\n\t"           \
  "s_nop 0\n\t"                                         \
  "v_cndmask_b32_e64 v2, %[plo], 0, vcc\n\t"            \
  "v_cndmask_b32_e64 v3, %[phi], 0, vcc\n\t"            \
  "v_lshl_add_u64 %[acc], v[2:3], 0, %[acc]"

__device__ __forceinline__ void mac_step_asm(uint64_t &acc, uint64_t x, uint64_t y, uint64_t yq,
                                             const StepConsts &k) {
  const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32), yl = (uint32_t)y, yh = (uint32_t)(y >> 32);
  const uint32_t ql = (uint32_t)yq, qh = (uint32_t)(yq >> 32);
  const uint64_t T = __umulhi(xl, ql);
  asm(XPG_MAC_STEP_ASM
      : [acc] "+v"(acc)
      : [xl] "v"(xl), [xh] "v"(xh), [yl] "v"(yl), [yh] "v"(yh), [ql] "v"(ql), [qh] "v"(qh), [T] "v"(T),
        [d] "s"(k.delta), [k29] "s"(k.k29), [negp] "s"(k.negp), [plo] "v"(k.plo), [phi] "v"(k.phi)
      : "v2", "v3", "v4", "v5", "v6", "v7", "v8", "v9", "vcc");
}

// mac_step_asm that also counts, in nsub, the steps that subtract p (acc + r >= p): the
// compare's lane mask is still in vcc after the select, and one add-with-carry adds it.
__device__ __forceinline__ void mac_step_count_asm(uint64_t &acc, uint32_t &nsub, uint64_t x, uint64_t y,
                                                   uint64_t yq, const StepConsts &k) {
  const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32), yl = (uint32_t)y, yh = (uint32_t)(y >> 32);
  const uint32_t ql = (uint32_t)yq, qh = (uint32_t)(yq >> 32);
  const uint64_t T = __umulhi(xl, ql);
  asm(XPG_MAC_STEP_ASM "\n\t"
      "v_addc_co_u32 %[ns], vcc, 0, %[ns], vcc"
      : [acc] "+v"(acc), [ns] "+v"(nsub)
      : [xl] "v"(xl), [xh] "v"(xh), [yl] "v"(yl), [yh] "v"(yh), [ql] "v"(ql), [qh] "v"(qh), [T] "v"(T),
        [d] "s"(k.delta), [k29] "s"(k.k29), [negp] "s"(k.negp), [plo] "v"(k.plo), [phi] "v"(k.phi)
      : "v2", "v3", "v4", "v5", "v6", "v7", "v8", "v9", "vcc");
}
#undef XPG_MAC_STEP_ASM

// The lazy Shoup product x*w - hi64(x*wq)*p (mod 2^64) for p = 2^61 - delta as 13 VALU
// instructions (equal to shoup_lazy(x, w, wq, p) for every x, w, wq < 2^64): q as in
// mac_step_asm; the four products reaching only the high word (xl*wh, xh*wl,
// hi(q)*delta, lo(q)*0xE0000000) are summed by chained multiply-adds, whose low word a
// 64-bit shift moves to the high half, and two multiply-adds add xl*wl and lo(q)*delta.
// Scratch v[4:9] and vcc; two wait states separate the carry's write of vcc from its
// read as a lane mask.
__device__ __forceinline__ uint64_t shoup_lazy_asm(uint64_t x, uint64_t w, uint64_t wq, uint32_t delta) {
  const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32), wl = (uint32_t)w, wh = (uint32_t)(w >> 32);
  const uint32_t ql = (uint32_t)wq, qh = (uint32_t)(wq >> 32);
  const uint64_t T = __umulhi(xl, ql);
  uint64_t r;
  asm("v_mad_u64_u32 v[4:5], vcc, %[xh], %[ql], %[T]\n\t"
      "v_mad_u64_u32 v[6:7], vcc, %[xl], %[qh], v[4:5]\n\t"
      "v_mov_b32 v8, v7\n\t"
      "s_nop 0\n\t"
      "v_cndmask_b32_e64 v9, 0, 1, vcc\n\t"
      "v_mad_u64_u32 v[4:5], vcc, %[xh], %[qh], v[8:9]\n\t"
      "v_mad_u64_u32 v[8:9], vcc, %[xl], %[wh], 0\n\t"
      "v_mad_u64_u32 v[8:9], vcc, %[xh], %[wl], v[8:9]\n\t"
      "v_mad_u64_u32 v[8:9], vcc, v5, %[d], v[8:9]\n\t"
      "v_mad_u64_u32 v[8:9], vcc, v4, %[k29], v[8:9]\n\t"
      "v_lshlrev_b64 v[8:9], 32, v[8:9]\n\t"
      "v_mad_u64_u32 %[r], vcc, %[xl], %[wl], v[8:9]\n\t"
      "v_mad_u64_u32 %[r], vcc, v4, %[d], %[r]"
      : [r] "=&v"(r)
      : [xl] "v"(xl), [xh] "v"(xh), [wl] "v"(wl), [wh] "v"(wh), [ql] "v"(ql), [qh] "v"(qh), [T] "v"(T),
        [d] "s"(delta), [k29] "s"(0xE0000000u)
      : "v4", "v5", "v6", "v7", "v8", "v9", "vcc");
  return r;
}

// (u + v) reduced once by m2 (m2 = 2p, negm2 = 2^64 - 2p): the sum, the sum minus m2, and a
// select on sum < m2 -- five VALU instructions.  Scratch v[2:3], v[10:11] and vcc.
__device__ __forceinline__ uint64_t add_mod2p_asm(uint64_t u, uint64_t v, uint64_t m2, uint64_t negm2) {
  uint32_t lo, hi;
  asm("v_lshl_add_u64 v[2:3], %[u], 0, %[v]\n\t"
      "v_lshl_add_u64 v[10:11], v[2:3], 0, %[negm2]\n\t"
      "v_cmp_gt_u64_e64 vcc, %[m2], v[2:3]\n\t"
      "s_nop 1\n\t"
      "v_cndmask_b32_e64 %[lo], v10, v2, vcc\n\t"
      "v_cndmask_b32_e64 %[hi], v11, v3, vcc"
      : [lo] "=v"(lo), [hi] "=v"(hi)
      : [u] "v"(u), [v] "v"(v), [m2] "s"(m2), [negm2] "s"(negm2)
      : "v2", "v3", "v10", "v11", "vcc");
  return ((uint64_t)hi << 32) | lo;
}

// u - v, plus m2 when u < v (the borrow selects) -- five VALU instructions.  Scratch
// v[2:3], v[10:11] and vcc.
__device__ __forceinline__ uint64_t sub_mod2p_asm(uint64_t u, uint64_t v, uint64_t m2) {
  uint32_t lo, hi;
  asm("v_sub_co_u32_e64 v2, vcc, %[ul], %[vl]\n\t"
      "s_nop 1\n\t"
      "v_subb_co_u32_e64 v3, vcc, %[uh], %[vh], vcc\n\t"
      "v_lshl_add_u64 v[10:11], v[2:3], 0, %[m2]\n\t"
      "s_nop 1\n\t"
      "v_cndmask_b32_e64 %[lo], v2, v10, vcc\n\t"
      "v_cndmask_b32_e64 %[hi], v3, v11, vcc"
      : [lo] "=v"(lo), [hi] "=v"(hi)
      : [ul] "v"((uint32_t)u), [uh] "v"((uint32_t)(u >> 32)), [vl] "v"((uint32_t)v), [vh] "v"((uint32_t)(v >> 32)),
        [m2] "s"(m2)
      : "v2", "v3", "v10", "v11", "vcc");
  return ((uint64_t)hi << 32) | lo;
}

// counter-based SplitMix64: word k of stream `seed` (the seeded generator of this
// project; stands in for the reference's /dev/urandom-keyed Salsa20 stream).
__host__ __device__ __forceinline__ uint64_t rand_word(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
