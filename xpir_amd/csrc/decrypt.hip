// decrypt.hip -- client-side decryption of reply ciphertexts on the GPU (test support and
// load generation: checking many replies without a CPU in the loop).
//
// Reference: NFLLWE::dec (crypto/NFLLWE.cpp:412-458) -> b - a*s in the NTT domain,
// NFLlib::invnttAndPowInvPhi (crypto/NFLlib.hpp:613-626: inverse NTT, then phi^-j and
// n^-1), NFLlib::poly2mpz (crypto/NFLlib.cpp:504-532: CRT lift of each coefficient to
// [0, q), centred), and the low A bits of each coefficient packed little-endian
// (serializeData64/32, NFLlib.cpp:412-487) -- the bytes PIRReplyExtraction_internal
// (pir/replyExtraction/PIRReplyExtraction_internal.cpp:41-106) hands to the next level.
// Only residues matter here (the message sits in the low A bits of the centred value),
// so the inverse transform is the forward DIF core with the inverse root between two
// bit-reversal permutations, and every value is reduced to [0, p) before the lift.
#include "internal.hpp"
#include "ntt_dif.hpp"

namespace {

struct DecArgs {
  ModSet ms;
  uint32_t n, logn, m, A;
  const uint64_t *iw, *iwq, *ipt, *iptq;
  const uint64_t *sk;      // m x n NTT-domain secret key
  const uint64_t *ct;      // n_ct ciphertexts a[m][n] || b[m][n]
  uint64_t *res;           // n_ct x m x n canonical residues
  uint64_t *low;           // n_ct x n x 2: low 128 bits of each centred coefficient
  uint8_t *out;            // n_ct x A*n/8 bytes
  uint64_t n_ct;
  // Garner constants: digit c = (r_c - sum_{j<c} dig_j * pre[c][j]) * inv[c] mod p_c,
  // pre[c][j] = p_0 ... p_{j-1} mod p_c; the value is sum_c dig_c * prod[c] (limbs), and
  // values above M/2 are centred by subtracting M = p_0 ... p_{m-1}.
  uint64_t pre[XPG_MAX_M][XPG_MAX_M], preq[XPG_MAX_M][XPG_MAX_M];
  uint64_t inv[XPG_MAX_M], invq[XPG_MAX_M];
  uint64_t prod[XPG_MAX_M][XPG_MAX_M];  // limbs of p_0 ... p_{c-1}
  uint64_t M[XPG_MAX_M + 1];            // limbs of M
};

// x mod p for any 64-bit x: x = h*2^61 + l = h*p + (l + h*delta), l + h*delta < 2p
__device__ __forceinline__ uint64_t reduce64(uint64_t x, uint64_t p, uint32_t delta) {
  const uint64_t v = (x & ((1ULL << 61) - 1)) + (x >> 61) * (uint64_t)delta;
  return v >= p ? v - p : v;
}

__device__ __forceinline__ uint32_t bitrev(uint32_t i, uint32_t logn) { return __brev(i) >> (32 - logn); }

// one workgroup per (ciphertext, modulus): b - a*s, inverse transform, phi^-j * n^-1
__global__ void __launch_bounds__(1024) dec_ntt_kernel(DecArgs a) {
  extern __shared__ uint64_t sh[];
  const uint32_t n = a.n, m = a.m;
  const uint64_t t = blockIdx.x / m;
  const uint32_t c = blockIdx.x % m;
  const uint64_t p = a.ms.p[c];
  const uint32_t delta = a.ms.delta[c];
  const uint64_t mn = (uint64_t)m * n;
  const uint64_t *ca = a.ct + t * 2 * mn + (uint64_t)c * n, *cb = ca + mn;
  const uint64_t *s = a.sk + (uint64_t)c * n;
  for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
    const uint64_t sj = s[j];
    const uint64_t as = reduce64(shoup_lazy_sf(ca[j], sj, shoup_quot(sj, p, delta), delta), p, delta);
    const uint64_t b = reduce64(cb[j], p, delta);
    sh[bitrev(j, a.logn)] = b >= as ? b - as : b + p - as;
  }
  __syncthreads();
  const uint64_t *w = a.iw + (uint64_t)c * n, *wq = a.iwq + (uint64_t)c * n;
  dif_radix2_layers(sh, n, a.logn, w, wq, p, delta);
  const uint64_t W = w[n - 4 + 1], Wq = wq[n - 4 + 1];
  for (uint32_t b = threadIdx.x; b < n / 4; b += blockDim.x) {
    uint64_t z[4];
    const uint32_t base = 4 * b;
    dif_last_two(z, sh[base], sh[base + 1], sh[base + 2], sh[base + 3], W, Wq, p, delta);
#pragma unroll
    for (int i = 0; i < 4; i++) sh[base + i] = z[i];
  }
  __syncthreads();
  const uint64_t *ip = a.ipt + (uint64_t)c * n, *ipq = a.iptq + (uint64_t)c * n;
  uint64_t *r = a.res + (t * m + c) * n;
  for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
    const uint64_t x = shoup_lazy_sf(sh[bitrev(j, a.logn)], ip[j], ipq[j], delta);
    r[j] = x >= p ? x - p : x;
  }
}

// one thread per (ciphertext, coefficient): Garner lift, centring, low 128 bits
__global__ void dec_crt_kernel(DecArgs a) {
  const uint64_t id = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (id >= a.n_ct * a.n) return;
  const uint64_t t = id / a.n, i = id % a.n;
  const uint32_t m = a.m;
  uint64_t dig[XPG_MAX_M];
  for (uint32_t c = 0; c < m; c++) {
    const uint64_t p = a.ms.p[c];
    const uint32_t delta = a.ms.delta[c];
    uint64_t acc = 0;
    for (uint32_t j = 0; j < c; j++) {
      uint64_t v = shoup_lazy_sf(reduce64(dig[j], p, delta), a.pre[c][j], a.preq[c][j], delta);
      v = v >= p ? v - p : v;
      acc += v;
      acc = acc >= p ? acc - p : acc;
    }
    const uint64_t r = a.res[(t * m + c) * a.n + i];
    const uint64_t diff = r >= acc ? r - acc : r + p - acc;
    const uint64_t d = shoup_lazy_sf(diff, a.inv[c], a.invq[c], delta);
    dig[c] = d >= p ? d - p : d;
  }
  // value = sum_c dig_c * prod[c], in m + 1 limbs
  uint64_t v[XPG_MAX_M + 1] = {0, 0, 0, 0, 0};
  for (uint32_t c = 0; c < m; c++) {
    uint64_t carry = 0;
    for (uint32_t l = 0; l <= c && l < XPG_MAX_M; l++) {
      const uint64_t lo = dig[c] * a.prod[c][l], hi = __umul64hi(dig[c], a.prod[c][l]);
      uint64_t s = v[l] + lo;
      uint64_t cy = s < lo ? 1 : 0;
      s += carry;
      cy += s < carry ? 1 : 0;
      v[l] = s;
      carry = hi + cy;
    }
    for (uint32_t l = c + 1; l <= m && carry; l++) {
      v[l] += carry;
      carry = v[l] < carry ? 1 : 0;
    }
  }
  // centred: if 2v > M then v - M (mod 2^128 suffices for the low bits)
  bool gt = false, eq = true;
  for (int l = (int)m; l >= 0; l--) {
    const uint64_t tw = (v[l] << 1) | (l > 0 ? v[l - 1] >> 63 : 0);
    if (eq && tw != a.M[l]) {
      gt = tw > a.M[l];
      eq = false;
    }
  }
  uint64_t lo = v[0], hi = v[1];
  if (gt) {
    const uint64_t nlo = lo - a.M[0];
    hi = hi - a.M[1] - (lo < a.M[0] ? 1 : 0);
    lo = nlo;
  }
  a.low[id * 2] = lo;
  a.low[id * 2 + 1] = hi;
}

// one thread per output word: bits [64w, 64w + 64) of a ciphertext's plaintext, gathered
// from the coefficients whose A-bit fields overlap them
__global__ void dec_pack_kernel(DecArgs a) {
  const uint64_t words = (uint64_t)a.A * a.n / 64;  // per ciphertext
  const uint64_t id = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (id >= a.n_ct * words) return;
  const uint64_t t = id / words, wi = id % words;
  const uint64_t bit0 = wi * 64, A = a.A;
  uint64_t out = 0;
  for (uint64_t i = bit0 / A; i * A < bit0 + 64 && i < a.n; i++) {
    const uint64_t lo = a.low[(t * a.n + i) * 2], hi = a.low[(t * a.n + i) * 2 + 1];
    // the coefficient's A low bits sit at bit i*A of the plaintext
    const int64_t sh = (int64_t)(i * A) - (int64_t)bit0;  // position inside this word
    uint64_t fl = lo, fh = hi;                          // field = low A bits
    if (A < 128) {
      if (A >= 64) fh &= (A == 64) ? 0 : ((1ULL << (A - 64)) - 1);
      else {
        fl &= (1ULL << A) - 1;
        fh = 0;
      }
    }
    if (sh >= 0) {
      out |= fl << sh;  // sh < 64
    } else {
      const uint64_t s = (uint64_t)(-sh);  // drop s low bits of the 128-bit field
      if (s < 64) out |= (fl >> s) | (s ? fh << (64 - s) : 0);
      else if (s < 128) out |= fh >> (s - 64);
    }
  }
  reinterpret_cast<uint64_t *>(a.out)[id] = out;
}

// host: a*b mod p and modular inverse (Fermat)
uint64_t hmul(uint64_t x, uint64_t y, uint64_t p) { return (uint64_t)(((unsigned __int128)x * y) % p); }
uint64_t hpow(uint64_t x, uint64_t e, uint64_t p) {
  uint64_t r = 1 % p;
  for (; e; e >>= 1, x = hmul(x, x, p))
    if (e & 1) r = hmul(r, x, p);
  return r;
}
uint64_t hquot(uint64_t y, uint64_t p) { return (uint64_t)(((unsigned __int128)y << 64) / p); }

}  // namespace

extern "C" int xpg_decrypt_dev(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *sk_dev, const uint64_t *ct_dev,
                               uint64_t n_ct, uint8_t *plain_dev) {
  if (!ctx || !prm || !sk_dev || !ct_dev || !plain_dev) return set_err(ctx, XPG_EINVAL, "null argument");
  if (prm->m < 1 || prm->m > XPG_MAX_M || prm->n < 64 || (prm->n & (prm->n - 1)))
    return set_err(ctx, XPG_EINVAL, "unsupported parameters");
  if (prm->A <= 0 || prm->A > 128) return set_err(ctx, XPG_EINVAL, "decryption supports 1..128 absorption bits");
  if (reinterpret_cast<uintptr_t>(plain_dev) & 7) return set_err(ctx, XPG_EINVAL, "plaintext buffer not 8-byte aligned");
  if (n_ct == 0) return XPG_OK;
  const DevTables *tb = get_tables(ctx, prm->n, prm->m);
  if (!tb) return XPG_EDEVICE;
  const uint64_t n = prm->n, m = prm->m;
  if (n_ct * m > 0x7fffffffULL || n_ct * n > 0x7fffffffULL * 256ULL) return set_err(ctx, XPG_EINVAL, "too many ciphertexts");
  DecArgs a{};
  a.ms = modset_of(prm);
  a.n = prm->n;
  a.logn = prm->logn;
  a.m = prm->m;
  a.A = (uint32_t)prm->A;
  a.iw = tb->iw;
  a.iwq = tb->iwq;
  a.ipt = tb->ipt;
  a.iptq = tb->iptq;
  a.sk = sk_dev;
  a.ct = ct_dev;
  a.out = plain_dev;
  a.n_ct = n_ct;
  for (uint32_t c = 0; c < m; c++) {
    const uint64_t p = prm->moduli[c];
    uint64_t pr = 1;
    for (uint32_t j = 0; j < c; j++) {
      a.pre[c][j] = pr;
      a.preq[c][j] = hquot(pr, p);
      pr = hmul(pr, prm->moduli[j] % p, p);
    }
    a.inv[c] = hpow(pr, p - 2, p);  // (p_0 ... p_{c-1})^-1 mod p_c
    a.invq[c] = hquot(a.inv[c], p);
    // limbs of p_0 ... p_{c-1}
    uint64_t limbs[XPG_MAX_M + 1] = {1, 0, 0, 0, 0};
    for (uint32_t j = 0; j < c; j++) {
      unsigned __int128 carry = 0;
      for (uint32_t l = 0; l <= XPG_MAX_M; l++) {
        const unsigned __int128 x = (unsigned __int128)limbs[l] * prm->moduli[j] + carry;
        limbs[l] = (uint64_t)x;
        carry = x >> 64;
      }
    }
    for (uint32_t l = 0; l < XPG_MAX_M; l++) a.prod[c][l] = limbs[l];
    if (c + 1 == m) {
      unsigned __int128 carry = 0;
      for (uint32_t l = 0; l <= XPG_MAX_M; l++) {
        const unsigned __int128 x = (unsigned __int128)limbs[l] * p + carry;
        a.M[l] = (uint64_t)x;
        carry = x >> 64;
      }
    }
  }
  // scratch: residues n_ct x m x n, then the low 128 bits n_ct x n x 2
  int rc = ensure_dec_buf(ctx, (n_ct * m * n + n_ct * n * 2) * 8);
  if (rc) return rc;
  a.res = ctx->dec_buf;
  a.low = ctx->dec_buf + n_ct * m * n;
  uint32_t threads = (uint32_t)(n / 4 > 1024 ? 1024 : n / 4);
  hipLaunchKernelGGL(dec_ntt_kernel, dim3((uint32_t)(n_ct * m)), dim3(threads), n * sizeof(uint64_t), ctx->stream, a);
  if (hipGetLastError() != hipSuccess) return set_err(ctx, XPG_EDEVICE, "decrypt transform launch");
  hipLaunchKernelGGL(dec_crt_kernel, dim3((uint32_t)((n_ct * n + 255) / 256)), dim3(256), 0, ctx->stream, a);
  if (hipGetLastError() != hipSuccess) return set_err(ctx, XPG_EDEVICE, "decrypt lift launch");
  const uint64_t words = n_ct * (uint64_t)a.A * n / 64;
  hipLaunchKernelGGL(dec_pack_kernel, dim3((uint32_t)((words + 255) / 256)), dim3(256), 0, ctx->stream, a);
  if (hipGetLastError() != hipSuccess) return set_err(ctx, XPG_EDEVICE, "decrypt pack launch");
  return XPG_OK;
}
