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
// twiddles); layers run out of LDS, two per barrier (ntt_dif.hpp), the last two layers
// fused in registers as the reference fuses them.
#include "internal.hpp"
#include "ntt_dif.hpp"

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

// bits [bitpos, bitpos+len) of unit u (len <= 64), zero past the unit's clipped end
__device__ __forceinline__ uint64_t unit_bits(const ByteSrc &s, uint64_t u, uint64_t bitpos, uint32_t len) {
  const uint64_t ustart = u * s.stride;
  uint64_t uend = ustart + s.len;
  if (uend > s.limit) uend = s.limit;
  if (uend <= ustart) return 0;
  const uint64_t ubits = (uend - ustart) * 8;
  if (bitpos >= ubits) return 0;
  if (bitpos + len > ubits) len = (uint32_t)(ubits - bitpos);
  const uint64_t abit = ustart * 8 + bitpos;
  const uint64_t aw = (abit >> 3) & ~7ULL;
  const uint32_t o = (uint32_t)(abit & 63);
  uint64_t v = load_word_clipped(s.base, aw, uend) >> o;
  if (o != 0 && o + len > 64) v |= load_word_clipped(s.base, aw + 8, uend) << (64 - o);
  return len >= 64 ? v : (v & ((1ULL << len) - 1));
}

// absorbed coefficient (chunk `chunk` of unit u) for modulus c: the raw A-bit chunk when
// A <= 56, else sum_t (sub_t * 2^(56t) mod p) left unreduced (NFLlib.cpp:884-916).
__device__ __forceinline__ uint64_t absorbed(const ByteSrc &s, const AbsorbConsts &ac, uint64_t p,
                                             uint32_t c, uint64_t u, uint64_t chunk) {
  const uint64_t bit0 = chunk * ac.A;
  if (ac.nsub == 1) return unit_bits(s, u, bit0, ac.A);
  uint64_t v = 0;
  for (uint32_t t = 0; t < ac.nsub; t++) {
    uint32_t len = (t + 1 < ac.nsub) ? 56 : ac.A - 56 * (ac.nsub - 1);
    uint64_t sub = unit_bits(s, u, bit0 + 56ULL * t, len);
    v += shoup_canon(sub, ac.mult[c][t], ac.multq[c][t], p);
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

template <int KIND>
__global__ void __launch_bounds__(1024) ntt_kernel(NttArgs a, ByteSrc bs, AbsorbConsts ac, EncSrc es,
                                                    uint64_t *__restrict__ out) {
  extern __shared__ uint64_t sh[];
  const uint32_t n = a.n, m = a.m;
  const uint64_t blk = blockIdx.x;
  const uint32_t c = (uint32_t)(blk % m);
  const uint64_t uk = blk / m;
  const uint32_t k = (uint32_t)(uk % a.P);
  const uint64_t u = uk / a.P;
  const uint64_t p = a.ms.p[c];
  const uint32_t delta = a.ms.delta[c];
  const uint64_t *phi = a.phi + (size_t)c * n, *phiq = a.phiq + (size_t)c * n;
  const uint64_t *w = a.w + (size_t)c * n, *wq = a.wq + (size_t)c * n;

  // load + absorb + phi twist (canonical, NFLlib.hpp:601-604)
  for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
    uint64_t v;
    if (KIND == SRC_BYTES) {
      v = absorbed(bs, ac, p, c, u, (uint64_t)k * n + j);
    } else {
      v = enc_pre(es, p, c, es.elem0 + u, n, j, u == es.hot ? 1u : 0u);
    }
    sh[j] = shoup_canon(v, phi[j], phiq[j], p);
  }
  __syncthreads();

  dif_radix2_layers(sh, n, a.logn, w, wq, p, delta);

  // last two layers fused (NFLlib.hpp:502-535), straight to HBM
  const uint64_t W = w[n - 4 + 1], Wq = wq[n - 4 + 1];
  for (uint32_t b = threadIdx.x; b < n / 4; b += blockDim.x) {
    const uint32_t base = 4 * b;
    uint64_t z[4];
    dif_last_two(z, sh[base], sh[base + 1], sh[base + 2], sh[base + 3], W, Wq, p, delta);
    if (KIND == SRC_BYTES) {
      uint64_t *o = out + (((u * a.P) + k) * m + c) * (uint64_t)n + base;
#pragma unroll
      for (int i = 0; i < 4; i++) o[i] = z[i];
    } else {
      // b = b + a*s (one lazy MAC step), then a, b reduced (NFLLWE.cpp:339-357)
      const uint64_t elem = es.elem0 + u, mn = (uint64_t)m * n;
      uint64_t *oa = out + u * 2 * mn + (uint64_t)c * n + base, *ob = oa + mn;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint32_t pos = base + i;
        const uint64_t av = uniform_mod(rand_word(es.noise_seed ^ kAStreamXor, elem * mn + (uint64_t)c * n + pos), p);
        const uint64_t s = es.sk[(uint64_t)c * n + pos];
        uint64_t bv = mac_step(z[i], av, s, shoup_quot(s, p, a.ms.delta[c]), p, a.ms.delta[c]);
        bv = bv >= p ? bv - p : bv;
        bv = bv >= p ? bv - p : bv;
        oa[i] = av;
        ob[i] = bv;
      }
    }
  }
}

__global__ void fill_zero_kernel(uint64_t *dst, uint64_t words) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = 0;
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

uint32_t block_threads(uint32_t n) {
  uint32_t t = n / 4;
  return t > 1024 ? 1024 : t;
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
    hipLaunchKernelGGL(ntt_kernel<SRC_BYTES>, dim3((uint32_t)blocks), dim3(block_threads(prm->n)),
                       prm->n * sizeof(uint64_t), ctx->stream, a, src, ac, es, out);
    if (hipGetLastError() != hipSuccess) return set_err(ctx, XPG_EDEVICE, "import launch");
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
  hipLaunchKernelGGL(ntt_kernel<SRC_ENCRYPT>, dim3((uint32_t)blocks), dim3(block_threads(prm->n)),
                     prm->n * sizeof(uint64_t), ctx->stream, a, bs, ac, es, out);
  return hipGetLastError() == hipSuccess ? XPG_OK : set_err(ctx, XPG_EDEVICE, "encrypt launch");
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
