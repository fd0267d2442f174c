// import.hip -- database import for CDNA4: absorb encoding + phi twist + Harvey NTT,
// fused in one kernel per (unit, poly, modulus) block of n words.
//
// Reference path: PIRReplyGeneratorNFL_internal::importDataNFL
// (pir/replyGenerator/PIRReplyGeneratorNFL_internal.cpp:79-158) -> NFLLWE::deserializeDataNFL
// (crypto/NFLLWE.cpp:146-148) -> NFLlib::deserializeDataNFL (crypto/NFLlib.cpp:375-401):
// bitsplitter (:846-916) then nttAndPowPhi (crypto/NFLlib.hpp:597-608) / ntt (:450-537).
// The recursion step fromResulttoInData (PIRReplyGeneratorNFL_internal.cpp:599-626) is the
// same transform over the level's reply words, so it uses the same kernel.
//
// Exactness: the NTT output words are lazy in [0,2p) and depend on the exact butterfly
// sequence, so every layer is the reference's radix-2 DIF butterfly (no re-associated
// twiddles); three layers run per register pass between LDS round trips, the last two
// layers fused in registers as the reference fuses them.  The query encryption
// (NFLLWE::enc, crypto/NFLLWE.cpp:276-372) runs the same transform on its noise words.
#include "internal.hpp"

namespace {

constexpr uint64_t kAStreamXor = 0xA0A0A0A0A0A0A0A0ULL;  // stream of the uniform `a` words

struct AbsorbConsts {
  uint32_t A, nsub;
  uint64_t mult[XPG_MAX_M][4];   // 2^(56 t) mod p
  uint64_t multq[XPG_MAX_M][4];  // Shoup quotients
};

struct NttArgs {
  ModSet ms;
  uint32_t n, logn, m, P;
  const uint64_t *phi, *phiq, *w, *wq;
};

// 64-bit word at byte offset `aw` (multiple of 8) of base, bytes at or past `limit` as 0.
__device__ __forceinline__ uint64_t load_word_clipped(const uint8_t *base, uint64_t aw, uint64_t limit) {
  if (aw + 8 <= limit) return *reinterpret_cast<const uint64_t *>(base + aw);
  uint64_t v = 0;
  for (uint64_t b = aw; b < limit; b++) v |= (uint64_t)base[b] << (8 * (b - aw));
  return v;
}

// The bits of one unit: bytes [start, end) of base (its clipped extent).
struct UnitView {
  const uint8_t *base;
  uint64_t start, end, bits;
};

__device__ __forceinline__ UnitView unit_view(const ByteSrc &s, uint64_t u) {
  UnitView v;
  v.base = s.base;
  v.start = u * s.stride;
  v.end = v.start + s.len < s.limit ? v.start + s.len : s.limit;
  v.bits = v.end > v.start ? (v.end - v.start) * 8 : 0;
  return v;
}

// bits [bitpos, bitpos+len) of the unit (len <= 64), zero past its clipped end.  Words are
// read 8-byte aligned relative to base; the clipped byte loop runs only for the last word
// or two of a unit.
__device__ __forceinline__ uint64_t unit_bits(const UnitView &uv, uint64_t bitpos, uint32_t len) {
  if (bitpos >= uv.bits) return 0;
  if (bitpos + len > uv.bits) len = (uint32_t)(uv.bits - bitpos);
  const uint64_t abit = uv.start * 8 + bitpos;
  const uint64_t aw = (abit >> 3) & ~7ULL;
  const uint32_t o = (uint32_t)(abit & 63);
  uint64_t v;
  if (aw + 16 <= uv.end) {
    const uint64_t *wp = reinterpret_cast<const uint64_t *>(uv.base + aw);
    const uint64_t w0 = wp[0], w1 = wp[1];
    v = o ? (w0 >> o) | (w1 << (64 - o)) : w0;
  } else {
    v = load_word_clipped(uv.base, aw, uv.end) >> o;
    if (o != 0 && o + len > 64) v |= load_word_clipped(uv.base, aw + 8, uv.end) << (64 - o);
  }
  return len >= 64 ? v : (v & ((1ULL << len) - 1));
}

// absorbed coefficient (chunk `chunk` of the unit) for modulus c: the raw A-bit chunk when
// A <= 56, else sum_t (sub_t * 2^(56t) mod p) left unreduced (NFLlib.cpp:884-916); the
// t = 0 term is sub_0 itself (2^0 mod p = 1 and sub_0 < 2^56 < p).
__device__ __forceinline__ uint64_t absorbed(const UnitView &uv, const AbsorbConsts &ac, uint64_t p, uint32_t delta,
                                             uint32_t c, uint64_t chunk) {
  const uint64_t bit0 = chunk * ac.A;
  if (ac.nsub == 1) return unit_bits(uv, bit0, ac.A);
  if (bit0 >= uv.bits) return 0;
  uint64_t v = unit_bits(uv, bit0, 56);
  for (uint32_t t = 1; t < ac.nsub; t++) {
    const uint32_t len = (t + 1 < ac.nsub) ? 56 : ac.A - 56 * (ac.nsub - 1);
    const uint64_t sub = unit_bits(uv, bit0 + 56ULL * t, len);
    const uint64_t r = shoup_lazy_asm(sub, ac.mult[c][t], ac.multq[c][t], delta);
    v += r >= p ? r - p : r;
  }
  return v;
}

// pre-NTT b coefficient of an encryption of `bit`: e*2^A mod p plus the message
// (crypto/NFLLWE.cpp:293-330); e is the bounded noise draw (NFLlib.cpp:286-322).
__device__ __forceinline__ uint64_t enc_pre(const EncSrc &s, uint64_t p, uint32_t c, uint64_t elem,
                                            uint32_t n, uint32_t j, uint32_t bit) {
  uint64_t e = rand_word(s.noise_seed, elem * n + j) & s.emask;
  if (e >= s.ebound) e -= s.ebound;
  uint64_t v = shoup_canon(e, s.amp[c], s.ampq[c], p) + (j == 0 ? bit : 0);
  return v >= p ? v - p : v;
}

__device__ __forceinline__ uint64_t uniform_mod(uint64_t r, uint64_t p) {
  r &= (1ULL << 61) - 1;  // ceil(lg p) = 61 bits for every modulus
  return r >= p ? r - p : r;
}

// ---- the transform: radix-2 DIF layers three per register pass ----
//
// One workgroup per (unit, poly, modulus) block of N = 2^LOGN words, T = N/8 threads.  A
// pass of K <= 3 consecutive layers L0 .. L0+K-1 takes groups of 2^K words that only
// exchange among themselves: with h0 = N >> (L0+1) the first layer's half block and
// E = h0 >> (K-1), group gi covers words g + s*E (s < 2^K), g = (gi / E) * 2*h0 + gi % E.
// A thread holds 8 words: groups gi = t + r*T for r < 8/2^K.  Layer L0 + l pairs words
// s and s + 2^(K-1-l) with twiddle w[N - 2h + j0 + (s mod 2^(K-1-l))*E], h = h0 >> l and
// j0 = gi % E the group's offset in its block -- every butterfly is the reference's
// (crypto/NFLlib.hpp:478-499) on the same inputs and twiddle, so the words are the same
// in any pass grouping.  Passes meet in LDS, which is padded by one word per 8 (word i at
// i + i/8) so that each pass's accesses -- consecutive words across lanes, groups of 8
// consecutive words per lane in the last pass -- spread over all banks.  The first pass
// (layers 0-2) runs on the absorbed words in registers; the last pass takes layer
// log2(N) - 3 on 8 consecutive words and then the reference's fused last two layers
// (:502-535), and writes to HBM.

__device__ __forceinline__ uint32_t lds_at(uint32_t i) { return i + (i >> 3); }

struct BflyK {
  uint64_t p2, negp2;  // 2p and 2^64 - 2p
  uint32_t delta;
};

__device__ __forceinline__ uint64_t add_mod2p(uint64_t u, uint64_t v, const BflyK &k) {
  return add_mod2p_asm(u, v, k.p2, k.negp2);
}

// the reference's lazy DIF butterfly on [0, 2p) words: (u, v) -> (u + v, (u - v) * w)
__device__ __forceinline__ void bfly(uint64_t &u, uint64_t &v, uint64_t w, uint64_t wq, const BflyK &k) {
  const uint64_t s = add_mod2p(u, v, k);
  v = shoup_lazy_asm(u - v + k.p2, w, wq, k.delta);
  u = s;
}

template <int LOGN, int L0, int K>
__device__ __forceinline__ uint32_t group_base(uint32_t t, int r) {
  constexpr uint32_t N = 1u << LOGN, T = N / 8, H0 = N >> (L0 + 1), E = H0 >> (K - 1);
  const uint32_t gi = t + r * T;
  return (gi / E) * 2 * H0 + (gi & (E - 1));
}

template <int LOGN, int L0, int K>
__device__ __forceinline__ void pass_layers(uint64_t (&x)[8], uint32_t t, const uint64_t *w, const uint64_t *wq,
                                            const BflyK &k) {
  constexpr uint32_t N = 1u << LOGN, T = N / 8, H0 = N >> (L0 + 1), E = H0 >> (K - 1);
  constexpr int G = 1 << K;
#pragma unroll
  for (int r = 0; r < 8 / G; r++) {
    const uint32_t j0 = (t + r * T) & (E - 1);
#pragma unroll
    for (int l = 0; l < K; l++) {
      const int dist = 1 << (K - 1 - l);
      const uint32_t off = N - 2 * (H0 >> l) + j0;
#pragma unroll
      for (int s = 0; s < G; s++) {
        if (s & dist) continue;
        const uint32_t ti = off + (uint32_t)(s & (dist - 1)) * E;
        bfly(x[r * G + s], x[r * G + s + dist], w[ti], wq[ti], k);
      }
    }
  }
}

template <int LOGN, int L0, int K>
__device__ __forceinline__ void pass_store(const uint64_t (&x)[8], uint64_t *sh, uint32_t t) {
  constexpr uint32_t N = 1u << LOGN, H0 = N >> (L0 + 1), E = H0 >> (K - 1);
  constexpr int G = 1 << K;
#pragma unroll
  for (int r = 0; r < 8 / G; r++) {
    const uint32_t g = group_base<LOGN, L0, K>(t, r);
#pragma unroll
    for (int s = 0; s < G; s++) sh[lds_at(g + s * E)] = x[r * G + s];
  }
}

template <int LOGN, int L0, int K>
__device__ __forceinline__ void pass_load(uint64_t (&x)[8], const uint64_t *sh, uint32_t t) {
  constexpr uint32_t N = 1u << LOGN, H0 = N >> (L0 + 1), E = H0 >> (K - 1);
  constexpr int G = 1 << K;
#pragma unroll
  for (int r = 0; r < 8 / G; r++) {
    const uint32_t g = group_base<LOGN, L0, K>(t, r);
#pragma unroll
    for (int s = 0; s < G; s++) x[r * G + s] = sh[lds_at(g + s * E)];
  }
}

// Passes from layer L (x holds the words in the grouping of the pass that ended at L,
// stored to LDS below) up to layer LOGN - 4; the last pass's layout is left in LDS.
template <int LOGN, int L, int KPREV>
__device__ __forceinline__ void middle_passes(uint64_t (&x)[8], uint64_t *sh, uint32_t t, const uint64_t *w,
                                              const uint64_t *wq, const BflyK &k) {
  pass_store<LOGN, L - KPREV, KPREV>(x, sh, t);
  __syncthreads();
  if constexpr (L < LOGN - 3) {
    constexpr int K = (LOGN - 3 - L) < 3 ? (LOGN - 3 - L) : 3;
    pass_load<LOGN, L, K>(x, sh, t);
    pass_layers<LOGN, L, K>(x, t, w, wq, k);
    __syncthreads();
    middle_passes<LOGN, L + K, K>(x, sh, t, w, wq, k);
  }
}

// the reference's fused last two layers on four consecutive words (NFLlib.hpp:502-535)
__device__ __forceinline__ void last_two(uint64_t &u0, uint64_t &u1, uint64_t &u2, uint64_t &u3, uint64_t W,
                                         uint64_t Wq, const BflyK &k) {
  const uint64_t p2 = k.p2;
  const uint64_t v0 = add_mod2p(u0, u2, k);
  const uint64_t v2 = sub_mod2p_asm(u0, u2, p2);
  const uint64_t v1 = add_mod2p(u1, u3, k);
  const uint64_t v3 = shoup_lazy_asm(u1 - u3 + p2, W, Wq, k.delta);
  u0 = add_mod2p(v0, v1, k);
  u1 = sub_mod2p_asm(v0, v1, p2);
  u2 = add_mod2p(v2, v3, k);
  u3 = sub_mod2p_asm(v2, v3, p2);
}

template <int KIND, int LOGN>
__global__ void __launch_bounds__((1 << LOGN) / 8) ntt8_kernel(NttArgs a, ByteSrc bs, AbsorbConsts ac, EncSrc es,
                                                                uint64_t *__restrict__ out) {
  constexpr uint32_t N = 1u << LOGN, T = N / 8;
  extern __shared__ uint64_t sh[];
  const uint32_t m = a.m, t = threadIdx.x;
  const uint64_t blk = blockIdx.x;
  const uint32_t c = (uint32_t)(blk % m);
  const uint64_t uk = blk / m;
  const uint32_t k = (uint32_t)(uk % a.P);
  const uint64_t u = uk / a.P;
  const uint64_t p = a.ms.p[c];
  const uint32_t delta = a.ms.delta[c];
  const BflyK bk{2 * p, 0 - 2 * p, delta};
  const uint64_t *phi = a.phi + (size_t)c * N, *phiq = a.phiq + (size_t)c * N;
  const uint64_t *w = a.w + (size_t)c * N, *wq = a.wq + (size_t)c * N;

  // absorb + phi twist (canonical, NFLlib.hpp:601-604) of the first pass's words t + s*T;
  // a word every lane of the wave absorbs as zero stays zero through the twist
  uint64_t x[8];
  const UnitView uv = unit_view(bs, u);
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const uint32_t j = t + s * T;
    uint64_t v;
    if (KIND == SRC_BYTES) {
      v = absorbed(uv, ac, p, delta, c, (uint64_t)k * N + j);
    } else {
      v = enc_pre(es, p, c, es.elem0 + u, N, j, u == es.hot ? 1u : 0u);
    }
    if (__any(v != 0)) {
      const uint64_t r = shoup_lazy_asm(v, phi[j], phiq[j], delta);
      v = r >= p ? r - p : r;
    }
    x[s] = v;
  }
  pass_layers<LOGN, 0, 3>(x, t, w, wq, bk);
  middle_passes<LOGN, 3, 3>(x, sh, t, w, wq, bk);

  // last pass: layer LOGN - 3 (pairs at distance 4) and the fused last two on the words
  // 8t .. 8t+7, straight to HBM
#pragma unroll
  for (int s = 0; s < 8; s++) x[s] = sh[lds_at(8 * t + s)];
#pragma unroll
  for (int s = 0; s < 4; s++) bfly(x[s], x[s + 4], w[N - 8 + s], wq[N - 8 + s], bk);
  const uint64_t W = w[N - 3], Wq = wq[N - 3];
  last_two(x[0], x[1], x[2], x[3], W, Wq, bk);
  last_two(x[4], x[5], x[6], x[7], W, Wq, bk);
  if (KIND == SRC_BYTES) {
    uint64_t *o = out + (((u * a.P) + k) * m + c) * (uint64_t)N + 8 * t;
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = x[i];
  } else {
    // b = b + a*s (one lazy MAC step), then a, b reduced (NFLLWE.cpp:339-357)
    const uint64_t elem = es.elem0 + u, mn = (uint64_t)m * N;
    uint64_t *oa = out + u * 2 * mn + (uint64_t)c * N + 8 * t, *ob = oa + mn;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const uint32_t pos = 8 * t + i;
      const uint64_t av = uniform_mod(rand_word(es.noise_seed ^ kAStreamXor, elem * mn + (uint64_t)c * N + pos), p);
      const uint64_t sv = es.sk[(uint64_t)c * N + pos];
      uint64_t bv = mac_step(x[i], av, sv, shoup_quot(sv, p, delta), p, delta);
      bv = bv >= p ? bv - p : bv;
      bv = bv >= p ? bv - p : bv;
      oa[i] = av;
      ob[i] = bv;
    }
  }
}

__global__ void fill_zero_kernel(uint64_t *dst, uint64_t words) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = 0;
}

// Order-keyed checksum of n words: the sum over i of SplitMix64's output for word i
// keyed by i (rand_word(w[i], i)), mod 2^64; one add per wave into *out (cleared before).
__global__ void word_checksum_kernel(const uint64_t *w, uint64_t n, unsigned long long *out) {
  uint64_t s = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    s += rand_word(w[i], i);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down((unsigned long long)s, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, (unsigned long long)s);
}

__global__ void random_bytes_kernel(uint64_t seed, uint64_t offset, uint8_t *dst, uint64_t nbytes) {
  // each thread writes the bytes of one stream word that overlap [offset, offset+nbytes)
  const uint64_t w0 = offset >> 3, w1 = (offset + nbytes + 7) >> 3;
  for (uint64_t wi = w0 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; wi < w1;
       wi += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = rand_word(seed, wi);
    const uint64_t b0 = wi * 8;
    if (b0 >= offset && b0 + 8 <= offset + nbytes && ((b0 - offset) & 7) == 0 &&
        ((reinterpret_cast<uintptr_t>(dst) & 7) == 0)) {
      *reinterpret_cast<uint64_t *>(dst + (b0 - offset)) = v;
    } else {
      for (int i = 0; i < 8; i++) {
        const uint64_t b = b0 + i;
        if (b >= offset && b < offset + nbytes) dst[b - offset] = (uint8_t)(v >> (8 * i));
      }
    }
  }
}

NttArgs make_args(const xpg_params *prm, const DevTables *t, uint32_t P) {
  NttArgs a;
  a.ms = modset_of(prm);
  a.n = prm->n;
  a.logn = prm->logn;
  a.m = prm->m;
  a.P = P;
  a.phi = t->phi;
  a.phiq = t->phiq;
  a.w = t->w;
  a.wq = t->wq;
  return a;
}

uint64_t host_mulmod(uint64_t a, uint64_t b, uint64_t p) {
  return (uint64_t)(((unsigned __int128)a * b) % p);
}
uint64_t host_quot(uint64_t y, uint64_t p) { return (uint64_t)(((unsigned __int128)y << 64) / p); }

// the transform kernel for n = 2^logn (512 .. 8192): N/8 threads, N + N/8 LDS words
template <int KIND>
int launch_ntt8(xpg_ctx *ctx, uint32_t logn, uint64_t blocks, const NttArgs &a, const ByteSrc &bs,
                const AbsorbConsts &ac, const EncSrc &es, uint64_t *out) {
  const uint32_t n = 1u << logn;
  const dim3 grid((uint32_t)blocks), wg(n / 8);
  const size_t lds = (n + n / 8) * sizeof(uint64_t);
  switch (logn) {
    case 9: hipLaunchKernelGGL((ntt8_kernel<KIND, 9>), grid, wg, lds, ctx->stream, a, bs, ac, es, out); break;
    case 10: hipLaunchKernelGGL((ntt8_kernel<KIND, 10>), grid, wg, lds, ctx->stream, a, bs, ac, es, out); break;
    case 11: hipLaunchKernelGGL((ntt8_kernel<KIND, 11>), grid, wg, lds, ctx->stream, a, bs, ac, es, out); break;
    case 12: hipLaunchKernelGGL((ntt8_kernel<KIND, 12>), grid, wg, lds, ctx->stream, a, bs, ac, es, out); break;
    case 13: hipLaunchKernelGGL((ntt8_kernel<KIND, 13>), grid, wg, lds, ctx->stream, a, bs, ac, es, out); break;
    default: return set_err(ctx, XPG_EINVAL, "transform degree outside 512 .. 8192");
  }
  return hipGetLastError() == hipSuccess ? XPG_OK : set_err(ctx, XPG_EDEVICE, "transform launch");
}

}  // namespace

int launch_fill_zero(xpg_ctx *ctx, uint64_t *dst, uint64_t words) {
  if (words == 0) return XPG_OK;
  uint64_t blocks = (words + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_zero_kernel, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, dst, words);
  return hipGetLastError() == hipSuccess ? XPG_OK : set_err(ctx, XPG_EDEVICE, "fill_zero launch");
}

int launch_import_bytes(xpg_ctx *ctx, const xpg_params *prm, const ByteSrc &src, uint64_t units,
                        uint64_t units_total, uint32_t P, uint64_t *out) {
  const DevTables *t = get_tables(ctx, prm->n, prm->m);
  if (!t) return XPG_EDEVICE;
  AbsorbConsts ac{};
  ac.A = (uint32_t)prm->A;
  ac.nsub = (ac.A + 55) / 56;
  if (ac.nsub > 4) return set_err(ctx, XPG_EINVAL, "absorption size above 224 bits");
  for (uint32_t c = 0; c < prm->m; c++) {
    uint64_t p = prm->moduli[c];
    ac.mult[c][0] = 1;
    for (uint32_t s = 1; s < ac.nsub; s++) ac.mult[c][s] = host_mulmod(ac.mult[c][s - 1], 1ULL << 56, p);
    for (uint32_t s = 0; s < ac.nsub; s++) ac.multq[c][s] = host_quot(ac.mult[c][s], p);
  }
  NttArgs a = make_args(prm, t, P);
  EncSrc es{};
  const uint64_t blocks = units * P * prm->m;
  if (blocks > 0) {
    if (blocks > 0x7fffffffULL) return set_err(ctx, XPG_EINVAL, "import grid too large");
    const int rc = launch_ntt8<SRC_BYTES>(ctx, prm->logn, blocks, a, src, ac, es, out);
    if (rc) return rc;
  }
  const uint64_t per_unit = (uint64_t)P * prm->m * prm->n;
  return launch_fill_zero(ctx, out + units * per_unit, (units_total - units) * per_unit);
}

int launch_encrypt(xpg_ctx *ctx, const xpg_params *prm, const EncSrc &src_in, uint64_t count,
                   uint64_t *out) {
  const DevTables *t = get_tables(ctx, prm->n, prm->m);
  if (!t) return XPG_EDEVICE;
  EncSrc es = src_in;
  const uint64_t bound = 2ULL * prm->Berr - 1;
  uint32_t lg = 0;
  while ((1ULL << lg) < bound) lg++;  // ceil(lg bound), NFLlib.cpp:296
  es.ebound = bound;
  es.emask = (1ULL << lg) - 1;
  for (uint32_t c = 0; c < prm->m; c++) {
    uint64_t p = prm->moduli[c], v = 1 % p;
    for (int i = 0; i < prm->A; i++) v = host_mulmod(v, 2, p);
    es.amp[c] = v;
    es.ampq[c] = host_quot(v, p);
  }
  NttArgs a = make_args(prm, t, 1);
  ByteSrc bs{};
  AbsorbConsts ac{};
  const uint64_t blocks = count * prm->m;
  if (blocks == 0) return XPG_OK;
  if (blocks > 0x7fffffffULL) return set_err(ctx, XPG_EINVAL, "query grid too large");
  return launch_ntt8<SRC_ENCRYPT>(ctx, prm->logn, blocks, a, bs, ac, es, out);
}

int launch_word_checksum(xpg_ctx *ctx, const uint64_t *words, uint64_t n, uint64_t *out_dev) {
  int rc = launch_fill_zero(ctx, out_dev, 1);
  if (rc || n == 0) return rc;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(word_checksum_kernel, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, words, n,
                     reinterpret_cast<unsigned long long *>(out_dev));
  return hipGetLastError() == hipSuccess ? XPG_OK : set_err(ctx, XPG_EDEVICE, "checksum launch");
}

int launch_random_bytes(xpg_ctx *ctx, uint64_t seed, uint64_t offset, uint8_t *dst, uint64_t nbytes) {
  if (nbytes == 0) return XPG_OK;
  uint64_t words = ((offset + nbytes + 7) >> 3) - (offset >> 3);
  uint64_t blocks = (words + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(random_bytes_kernel, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, seed, offset,
                     dst, nbytes);
  return hipGetLastError() == hipSuccess ? XPG_OK : set_err(ctx, XPG_EDEVICE, "random bytes launch");
}
