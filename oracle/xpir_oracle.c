/*
 * xpir_oracle.c -- CPU restatement of XPIR's NFLLWE reply path (TEST INFRASTRUCTURE).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this file.
 * Declarations and the reference lines each function follows are in xpir_oracle.h.
 *
 * Parity unpinned: xo_encrypt_bit (query generation).  The reference's NFLLWE::enc needs
 * Boost and keys its PRNG from /dev/urandom, so it cannot produce comparison bytes; the
 * seeded query streams are this project's own inputs, checked only by decryption round
 * trips (decrypt(encrypt(bit)) == bit, decrypt(reply) == the chosen element).  Every
 * other function is pinned against the reference's own code (oracle/_ref), its golden
 * vectors, or a closed form (tests/test_oracle.py).
 */
#include "xpir_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* Moduli p_i = 2^61 - c_i*2^14 + 1 and the primitive 2^14-th roots of unity the
 * reference fixes for them (crypto/NFLParams.cpp:34,41).  The roots determine the
 * NTT-domain words, so they are data that has to be identical. */
static const uint64_t kP[17] = {
  2305843009213317121ULL, 2305843009213120513ULL, 2305843009212694529ULL,
  2305843009212399617ULL, 2305843009211662337ULL, 2305843009211596801ULL,
  2305843009211400193ULL, 2305843009210580993ULL, 2305843009210515457ULL,
  2305843009210023937ULL, 2305843009209057281ULL, 2305843009208795137ULL,
  2305843009208713217ULL, 2305843009208123393ULL, 2305843009207468033ULL,
  2305843009206976513ULL, 2305843009206845441ULL};
static const uint64_t kRoot[17] = {
  1187132827279672845ULL, 753478288701480417ULL, 717492273700781398ULL,
  1965831349000139179ULL, 2188074825493578002ULL, 232964966911499637ULL,
  395215708149128273ULL, 241993652537061162ULL, 370764889455808762ULL,
  1598724739621481826ULL, 610870591815427803ULL, 1880981984715338041ULL,
  1389374064030612248ULL, 1508488901479281364ULL, 2111887543055470169ULL,
  2020509703903079203ULL, 1120240637471598814ULL};

uint64_t xo_modulus(uint32_t i) { return i < 17 ? kP[i] : 0; }
uint64_t xo_root(uint32_t i) { return i < 17 ? kRoot[i] : 0; }

uint64_t xo_mulmod(uint64_t x, uint64_t y, uint64_t p) { return (uint64_t)(((u128)x * y) % p); }

static uint64_t powmod(uint64_t b, uint64_t e, uint64_t p) {
  uint64_t r = 1 % p;
  b %= p;
  while (e) {
    if (e & 1) r = xo_mulmod(r, b, p);
    b = xo_mulmod(b, b, p);
    e >>= 1;
  }
  return r;
}

/* 8192^-1 mod p (the reference tabulates it, crypto/NFLParams.cpp:45); Fermat. */
uint64_t xo_inv8192(uint32_t i) { return i < 17 ? powmod(8192, kP[i] - 2, kP[i]) : 0; }

uint64_t xo_shoup_quot(uint64_t y, uint64_t p) { return (uint64_t)(((u128)y << 64) / p); }

/* x*y - floor(x*yq/2^64)*p computed modulo 2^64: in [0,2p) for p < 2^62. */
static inline uint64_t shoup_lazy(uint64_t x, uint64_t y, uint64_t yq, uint64_t p) {
  uint64_t q = (uint64_t)(((u128)x * yq) >> 64);
  return x * y - q * p;
}

uint64_t xo_mulmod_shoup(uint64_t x, uint64_t y, uint64_t yq, uint64_t p) {
  uint64_t r = shoup_lazy(x, y, yq, p);
  return r >= p ? r - p : r;
}

/* one step of the reply chain: acc + lazy product, one conditional subtraction of p */
uint64_t xo_mac_step(uint64_t acc, uint64_t x, uint64_t y, uint64_t yq, uint64_t p) {
  acc += shoup_lazy(x, y, yq, p);
  return acc >= p ? acc - p : acc;
}

static uint32_t ilog2u(uint64_t v) {
  uint32_t l = 0;
  while ((1ULL << (l + 1)) <= v) l++;
  return l;
}

int xo_params_init(const char *desc, uint64_t n0, xo_params *out) {
  unsigned k = 0, n = 0, bits = 0;
  char name[8] = {0};
  memset(out, 0, sizeof(*out));
  if (!desc || sscanf(desc, "%3[A-Z]:%u:%u:%u", name, &k, &n, &bits) != 4) return -1;
  if (strcmp(name, "LWE") != 0) return -2;
  if (bits == 0 || bits % 60 != 0) return -3;               /* NFLlib.cpp:219-224 */
  if (n < 512 || n > 8192 || (n & (n - 1))) return -4;       /* kMin/kMaxPolyDegree */
  if (bits / 60 > XO_MAX_M) return -5;
  out->k = k;
  out->n = n;
  out->logn = ilog2u(n);
  out->m = bits / 60;
  out->bits = bits;
  out->Berr = 5 * k / 2;                                     /* NFLLWE.cpp:84 */
  for (uint32_t i = 0; i < out->m; i++) out->p[i] = kP[i];
  if (n0 == 0) n0 = 1;
  /* NFLLWE.cpp:656-668: evaluated in double precision, floor((q-1 - lg n0 - lg B - lg n)/2) */
  double a = floor(((double)(out->m * 60) - 1.0 - log2((double)n0) - log2((double)out->Berr) -
                    log2((double)n)) / 2.0);
  out->A = (int32_t)a;
  if (out->A <= 0 || out->A > 64 * 3) return -6;
  return 0;
}

/* layered twiddle table: layer sizes n/2, n/4, ..., 1 holding w^(i*2^L) (n-1 words) */
static void twiddle_layers(uint64_t *tab, uint64_t *tabq, uint64_t w, uint32_t n, uint64_t p) {
  for (uint32_t size = n; size >= 2; size >>= 1) {
    uint64_t wi = 1;
    for (uint32_t i = 0; i < size / 2; i++) {
      *tab++ = wi;
      *tabq++ = xo_shoup_quot(wi, p);
      wi = xo_mulmod(wi, w, p);
    }
    w = xo_mulmod(w, w, p);
  }
}

int xo_tables_init(const xo_params *prm, xo_tables *t) {
  uint32_t n = prm->n, m = prm->m;
  size_t words = (size_t)n * m;
  memset(t, 0, sizeof(*t));
  t->n = n;
  t->m = m;
  uint64_t **arrs[8] = {&t->phi, &t->phiq, &t->w, &t->wq, &t->iw, &t->iwq, &t->ipt, &t->iptq};
  for (int i = 0; i < 8; i++) {
    *arrs[i] = (uint64_t *)calloc(words, sizeof(uint64_t));
    if (!*arrs[i]) return -1;
  }
  for (uint32_t c = 0; c < m; c++) {
    uint64_t p = prm->p[c];
    /* phi: primitive 2n-th root = root^(2^(13 - log2 n)) (NFLlib.cpp:86-90) */
    uint64_t phi = kRoot[c];
    for (uint32_t i = 0; i < 13 - prm->logn; i++) phi = xo_mulmod(phi, phi, p);
    uint64_t v = 1;
    for (uint32_t j = 0; j < n; j++) {
      t->phi[c * n + j] = v;
      t->phiq[c * n + j] = xo_shoup_quot(v, p);
      v = xo_mulmod(v, phi, p);
    }
    /* v = phi^n = -1, so phi^-1 = phi^n * phi^(n-1) */
    uint64_t iphi = xo_mulmod(v, t->phi[c * n + n - 1], p);
    uint64_t invn = xo_mulmod(xo_inv8192(c), 8192 / n, p);
    v = invn;
    for (uint32_t j = 0; j < n; j++) {
      t->ipt[c * n + j] = v;
      t->iptq[c * n + j] = xo_shoup_quot(v, p);
      v = xo_mulmod(v, iphi, p);
    }
    twiddle_layers(t->w + c * n, t->wq + c * n, xo_mulmod(phi, phi, p), n, p);
    twiddle_layers(t->iw + c * n, t->iwq + c * n, xo_mulmod(iphi, iphi, p), n, p);
  }
  return 0;
}

void xo_tables_free(xo_tables *t) {
  free(t->phi); free(t->phiq); free(t->w); free(t->wq);
  free(t->iw); free(t->iwq); free(t->ipt); free(t->iptq);
  memset(t, 0, sizeof(*t));
}

/* Radix-2 DIF layers with Harvey's lazy butterflies, then a fused final radix-4 block.
 * Inputs below 2p (the first layer gets canonical words), outputs in [0,2p). */
void xo_ntt(uint64_t *x, const uint64_t *w, const uint64_t *wq, uint32_t n, uint64_t p) {
  const uint64_t p2 = 2 * p;
  uint32_t logn = ilog2u(n);
  for (uint32_t L = 0; L + 2 < logn; L++) {
    uint32_t half = n >> (L + 1);
    for (uint32_t base = 0; base < n; base += 2 * half) {
      for (uint32_t j = 0; j < half; j++) {
        uint64_t u0 = x[base + j], u1 = x[base + j + half];
        uint64_t s = u0 + u1;
        if (s >= p2) s -= p2;
        uint64_t d = u0 - u1 + p2;
        x[base + j] = s;
        x[base + j + half] = shoup_lazy(d, w[j], wq[j], p);
      }
    }
    w += half;
    wq += half;
  }
  /* the remaining two layers: table layer of size 2 is {1, w^(n/4)} */
  const uint64_t W = w[1], Wq = wq[1];
  for (uint32_t base = 0; base < n; base += 4) {
    uint64_t u0 = x[base], u1 = x[base + 1], u2 = x[base + 2], u3 = x[base + 3];
    uint64_t v0 = u0 + u2; if (v0 >= p2) v0 -= p2;
    uint64_t v2 = u0 - u2; if ((int64_t)v2 < 0) v2 += p2;
    uint64_t v1 = u1 + u3; if (v1 >= p2) v1 -= p2;
    uint64_t v3 = shoup_lazy(u1 - u3 + p2, W, Wq, p);
    uint64_t z0 = v0 + v1; if (z0 >= p2) z0 -= p2;
    uint64_t z1 = v0 - v1; if ((int64_t)z1 < 0) z1 += p2;
    uint64_t z2 = v2 + v3; if (z2 >= p2) z2 -= p2;
    uint64_t z3 = v2 - v3; if ((int64_t)z3 < 0) z3 += p2;
    x[base] = z0; x[base + 1] = z1; x[base + 2] = z2; x[base + 3] = z3;
  }
}

void xo_ntt_powphi(const xo_tables *t, const xo_params *prm, uint64_t *poly) {
  uint32_t n = prm->n;
  for (uint32_t c = 0; c < prm->m; c++) {
    uint64_t p = prm->p[c];
    uint64_t *x = poly + (size_t)c * n;
    for (uint32_t j = 0; j < n; j++)
      x[j] = xo_mulmod_shoup(x[j], t->phi[c * n + j], t->phiq[c * n + j], p);
    xo_ntt(x, t->w + (size_t)c * n, t->wq + (size_t)c * n, n, p);
  }
}

static uint32_t bitrev(uint32_t v, uint32_t bits) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < bits; i++) { r = (r << 1) | (v & 1); v >>= 1; }
  return r;
}

void xo_invntt_powinvphi(const xo_tables *t, const xo_params *prm, uint64_t *poly) {
  uint32_t n = prm->n;
  uint64_t *y = (uint64_t *)malloc(sizeof(uint64_t) * n);
  for (uint32_t c = 0; c < prm->m; c++) {
    uint64_t p = prm->p[c];
    uint64_t *x = poly + (size_t)c * n;
    for (uint32_t i = 0; i < n; i++) y[bitrev(i, prm->logn)] = x[i];
    xo_ntt(y, t->iw + (size_t)c * n, t->iwq + (size_t)c * n, n, p);
    for (uint32_t i = 0; i < n; i++) x[bitrev(i, prm->logn)] = y[i];
    for (uint32_t j = 0; j < n; j++)
      x[j] = xo_mulmod_shoup(x[j], t->ipt[c * n + j], t->iptq[c * n + j], p);
  }
  free(y);
}

/* ------------------------------ absorb ------------------------------ */

uint64_t xo_absorb_polys(uint64_t nbytes, uint32_t A, uint32_t n) {
  /* NFLlib.cpp:378 evaluates ceil() in double; exact for these magnitudes */
  uint64_t bits = nbytes * 8, per = (uint64_t)A * n;
  return (bits + per - 1) / per;
}

/* bits [pos, pos+len) of the little-endian bit string, zero past nbytes (len <= 64) */
static uint64_t get_bits(const uint8_t *d, uint64_t nbytes, uint64_t pos, uint32_t len) {
  uint64_t v = 0;
  for (uint32_t got = 0; got < len;) {
    uint64_t byte = (pos + got) >> 3;
    uint32_t off = (uint32_t)((pos + got) & 7);
    uint32_t take = 8 - off;
    if (take > len - got) take = len - got;
    uint64_t b = byte < nbytes ? d[byte] : 0;
    v |= ((b >> off) & ((1ULL << take) - 1)) << got;
    got += take;
  }
  return v;
}

/* Chunk j = bits [jA, jA+A) goes to slot j%n of poly j/n.  A <= 56: the raw chunk in
 * every modulus.  A > 56: 56-bit sub-chunks t weighted by (2^(56t) mod p), each term
 * reduced, their sum left unreduced (NFLlib.cpp:884-916). */
void xo_absorb(const uint8_t *data, uint64_t nbytes, const xo_params *prm, uint64_t *out) {
  const uint32_t n = prm->n, m = prm->m, A = (uint32_t)prm->A;
  const uint64_t P = xo_absorb_polys(nbytes, A, n);
  const uint64_t nchunks = (nbytes * 8 + A - 1) / A;
  const uint32_t nsub = (A + 55) / 56;
  uint64_t mult[XO_MAX_M][4];
  for (uint32_t c = 0; c < m; c++) {
    mult[c][0] = 1;
    for (uint32_t s = 1; s < nsub; s++) mult[c][s] = xo_mulmod(mult[c][s - 1], 1ULL << 56, prm->p[c]);
  }
  memset(out, 0, sizeof(uint64_t) * P * m * n);
  for (uint64_t j = 0; j < nchunks; j++) {
    uint64_t poly = j / n, slot = j % n;
    for (uint32_t c = 0; c < m; c++) {
      uint64_t v = 0;
      if (nsub == 1) {
        v = get_bits(data, nbytes, j * A, A);
      } else {
        for (uint32_t s = 0; s < nsub; s++) {
          uint32_t len = (s + 1 < nsub) ? 56 : A - 56 * (nsub - 1);
          uint64_t sub = get_bits(data, nbytes, j * A + 56ULL * s, len);
          v += xo_mulmod(sub, mult[c][s], prm->p[c]);
        }
      }
      out[(poly * m + c) * n + slot] = v;
    }
  }
}

void xo_import(const xo_tables *t, const xo_params *prm, const uint8_t *data, uint64_t nbytes,
               uint64_t *out) {
  uint64_t P = xo_absorb_polys(nbytes, (uint32_t)prm->A, prm->n);
  xo_absorb(data, nbytes, prm, out);
  for (uint64_t k = 0; k < P; k++) xo_ntt_powphi(t, prm, out + k * prm->m * prm->n);
}

/* ------------------------------ reply MAC ------------------------------ */

void xo_reply_level(const xo_params *prm, const uint64_t *db, const uint64_t *q,
                    uint64_t n_l, uint64_t groups, uint32_t P, uint64_t *out, int threads) {
  const uint64_t n = prm->n, m = prm->m, mn = m * n;
  /* Shoup quotients of the query, as the server precomputes them on pushQuery
   * (PIRReplyGeneratorNFL_internal.cpp:754-771). */
  uint64_t *qq = (uint64_t *)malloc(sizeof(uint64_t) * n_l * 2 * mn);
  for (uint64_t e = 0; e < n_l; e++)
    for (uint64_t h = 0; h < 2; h++)
      for (uint64_t c = 0; c < m; c++)
        for (uint64_t s = 0; s < n; s++) {
          uint64_t idx = (e * 2 + h) * mn + c * n + s;
          qq[idx] = xo_shoup_quot(q[idx], prm->p[c]);
        }
  (void)threads;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) collapse(2)
#endif
  for (uint64_t g = 0; g < groups; g++) {
    for (uint64_t k = 0; k < P; k++) {
      uint64_t *acc = out + (g * P + k) * 2 * mn; /* a block then b block */
      memset(acc, 0, sizeof(uint64_t) * 2 * mn);
      for (uint64_t e = 0; e < n_l; e++) {
        const uint64_t *x = db + ((g * n_l + e) * P + k) * mn;
        const uint64_t *ya = q + (e * 2) * mn, *yb = ya + mn;
        const uint64_t *yaq = qq + (e * 2) * mn, *ybq = yaq + mn;
        for (uint64_t c = 0; c < m; c++) {
          uint64_t p = prm->p[c];
          for (uint64_t s = 0; s < n; s++) {
            uint64_t i = c * n + s;
            acc[i] = xo_mac_step(acc[i], x[i], ya[i], yaq[i], p);
            acc[mn + i] = xo_mac_step(acc[mn + i], x[i], yb[i], ybq[i], p);
          }
        }
      }
    }
  }
  free(qq);
}

/* The same chains as xo_reply_level, with the threads split over slot blocks instead of
 * over (group, poly): each chain still runs its elements in ascending order on one
 * thread, so the words are the same; only the work split differs (all host cores busy
 * even when P = 1). */
void xo_reply_level_chains(const xo_params *prm, const uint64_t *db, const uint64_t *q,
                           uint64_t n_l, uint64_t groups, uint32_t P, uint64_t *out,
                           int threads, double *mac_seconds) {
  const uint64_t n = prm->n, m = prm->m, mn = m * n;
  const uint64_t SB = n < 256 ? n : 256; /* slots per work item (n is a power of two) */
  const uint64_t nsb = n / SB;
  uint64_t *qq = (uint64_t *)malloc(sizeof(uint64_t) * n_l * 2 * mn);
  (void)threads;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
  for (uint64_t e = 0; e < n_l; e++)
    for (uint64_t h = 0; h < 2; h++)
      for (uint64_t i = 0; i < mn; i++) {
        uint64_t idx = (e * 2 + h) * mn + i;
        qq[idx] = xo_shoup_quot(q[idx], prm->p[i / n]);
      }
#ifdef _OPENMP
  double t0 = omp_get_wtime();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) collapse(4)
#endif
  for (uint64_t g = 0; g < groups; g++) {
    for (uint64_t k = 0; k < P; k++) {
      for (uint64_t c = 0; c < m; c++) {
        for (uint64_t b = 0; b < nsb; b++) {
          const uint64_t p = prm->p[c], lo = c * n + b * SB;
          uint64_t *ra = out + (g * P + k) * 2 * mn + lo, *rb = ra + mn;
          for (uint64_t s = 0; s < SB; s++) ra[s] = rb[s] = 0;
          for (uint64_t e = 0; e < n_l; e++) {
            const uint64_t *x = db + ((g * n_l + e) * P + k) * mn + lo;
            const uint64_t *ya = q + (e * 2) * mn + lo, *yb = ya + mn;
            const uint64_t *yaq = qq + (e * 2) * mn + lo, *ybq = yaq + mn;
            for (uint64_t s = 0; s < SB; s++) ra[s] = xo_mac_step(ra[s], x[s], ya[s], yaq[s], p);
            for (uint64_t s = 0; s < SB; s++) rb[s] = xo_mac_step(rb[s], x[s], yb[s], ybq[s], p);
          }
        }
      }
    }
  }
#ifdef _OPENMP
  if (mac_seconds) *mac_seconds = omp_get_wtime() - t0;
#else
  if (mac_seconds) *mac_seconds = 0;
#endif
  free(qq);
}

uint32_t xo_next_polys(const xo_params *prm, uint32_t P_l) {
  uint64_t ct_bytes = 16ULL * prm->n * prm->m;
  uint64_t abs_bytes = (uint64_t)prm->A * prm->n / 8;
  return (uint32_t)((P_l * ct_bytes + abs_bytes - 1) / abs_bytes);
}

uint32_t xo_recursion(const xo_tables *t, const xo_params *prm, const uint64_t *reply,
                      uint64_t groups, uint32_t P_l, uint64_t *out) {
  const uint64_t mn = (uint64_t)prm->m * prm->n;
  const uint64_t bytes = (uint64_t)P_l * 2 * mn * 8;
  const uint32_t P2 = xo_next_polys(prm, P_l);
  for (uint64_t g = 0; g < groups; g++)
    xo_import(t, prm, (const uint8_t *)(reply + g * P_l * 2 * mn), bytes, out + g * P2 * mn);
  return P2;
}

/* ------------------------------ seeded stream ------------------------------ */

uint64_t xo_rand_word(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void xo_rand_bytes(uint64_t seed, uint64_t off, uint8_t *out, uint64_t nbytes) {
  for (uint64_t i = 0; i < nbytes; i++) {
    uint64_t b = off + i;
    out[i] = (uint8_t)(xo_rand_word(seed, b >> 3) >> (8 * (b & 7)));
  }
}

/* ------------------------------ client side (tests) ------------------------------ */

/* uniform draw as setBoundedRandomPoly(uniform) does it: mask to ceil(lg p) bits, one -p */
static uint64_t uniform_mod(uint64_t r, uint64_t p) {
  uint64_t mask = (1ULL << (ilog2u(p) + 1)) - 1;
  r &= mask;
  return r >= p ? r - p : r;
}

void xo_keygen(const xo_params *prm, uint64_t key_seed, uint64_t *sk) {
  const uint64_t n = prm->n;
  for (uint64_t c = 0; c < prm->m; c++)
    for (uint64_t i = 0; i < n; i++)
      sk[c * n + i] = uniform_mod(xo_rand_word(key_seed, c * n + i), prm->p[c]);
}

/* Stream layout of one query element `elem` under `noise_seed`:
 *   noise e_i   = word (elem*n + i)        of stream noise_seed
 *   uniform a   = word (elem*m*n + c*n + i) of stream noise_seed ^ 0xA0A0...  */
#define XO_A_STREAM_XOR 0xA0A0A0A0A0A0A0A0ULL

void xo_encrypt_bit(const xo_tables *t, const xo_params *prm, const uint64_t *sk,
                    uint32_t bit, uint64_t noise_seed, uint64_t elem, uint64_t *out) {
  const uint64_t n = prm->n, m = prm->m, mn = m * n;
  uint64_t *a = out, *b = out + mn;
  const uint64_t bound = 2ULL * prm->Berr - 1;
  const uint64_t emask = (1ULL << (ilog2u(bound - 1) + 1)) - 1; /* ceil(lg bound) bits */
  for (uint64_t i = 0; i < n; i++) {
    uint64_t e = xo_rand_word(noise_seed, elem * n + i) & emask;
    if (e >= bound) e -= bound;
    for (uint64_t c = 0; c < m; c++) b[c * n + i] = e;
  }
  for (uint64_t c = 0; c < m; c++) {
    uint64_t p = prm->p[c];
    uint64_t amp = powmod(2, (uint64_t)prm->A, p), ampq = xo_shoup_quot(amp, p);
    for (uint64_t i = 0; i < n; i++) {
      uint64_t v = xo_mulmod_shoup(b[c * n + i], amp, ampq, p);
      uint64_t msg = (i == 0) ? bit : 0;
      v += msg;
      if (v >= p) v -= p;
      b[c * n + i] = v;
    }
  }
  xo_ntt_powphi(t, prm, b);
  for (uint64_t c = 0; c < m; c++) {
    uint64_t p = prm->p[c];
    for (uint64_t i = 0; i < n; i++) {
      uint64_t av = uniform_mod(xo_rand_word(noise_seed ^ XO_A_STREAM_XOR, elem * mn + c * n + i), p);
      uint64_t s = sk[c * n + i];
      uint64_t bv = xo_mac_step(b[c * n + i], av, s, xo_shoup_quot(s, p), p);
      a[c * n + i] = av % p;
      b[c * n + i] = bv % p;
    }
  }
}

/* ---- small fixed-width unsigned integers for the CRT lift (m <= 4 -> < 2^244) ---- */
typedef struct { uint64_t l[5]; } big;
static void big_set(big *r, uint64_t v) { memset(r, 0, sizeof(*r)); r->l[0] = v; }
static void big_mul_small(big *r, uint64_t v) {
  u128 carry = 0;
  for (int i = 0; i < 5; i++) {
    u128 t = (u128)r->l[i] * v + carry;
    r->l[i] = (uint64_t)t;
    carry = t >> 64;
  }
}
static void big_add(big *r, const big *a) {
  u128 carry = 0;
  for (int i = 0; i < 5; i++) {
    u128 t = (u128)r->l[i] + a->l[i] + carry;
    r->l[i] = (uint64_t)t;
    carry = t >> 64;
  }
}
static int big_cmp(const big *a, const big *b) {
  for (int i = 4; i >= 0; i--) {
    if (a->l[i] != b->l[i]) return a->l[i] > b->l[i] ? 1 : -1;
  }
  return 0;
}

int xo_decrypt(const xo_tables *t, const xo_params *prm, const uint64_t *sk,
               const uint64_t *ct, uint8_t *out) {
  const uint64_t n = prm->n, m = prm->m, mn = m * n;
  const uint32_t A = (uint32_t)prm->A;
  if (A > 128) return -1;
  uint64_t *v = (uint64_t *)malloc(sizeof(uint64_t) * mn);
  for (uint64_t c = 0; c < m; c++) {
    uint64_t p = prm->p[c];
    for (uint64_t i = 0; i < n; i++) {
      uint64_t s = sk[c * n + i];
      uint64_t as = xo_mac_step(0, ct[c * n + i], s, xo_shoup_quot(s, p), p);
      uint64_t y = as < p ? as : as % p;
      uint64_t r = ct[mn + c * n + i] + (p - y);
      v[c * n + i] = r >= p ? r - p : r;
    }
  }
  xo_invntt_powinvphi(t, prm, v);
  /* Garner: value = c0 + c1 p0 + c2 p0 p1 + ...; centred lift, keep the low A bits */
  big M;
  big_set(&M, 1);
  for (uint64_t c = 0; c < m; c++) big_mul_small(&M, prm->p[c]);
  memset(out, 0, A * n / 8);
  for (uint64_t i = 0; i < n; i++) {
    uint64_t dig[XO_MAX_M];
    for (uint64_t c = 0; c < m; c++) {
      uint64_t p = prm->p[c];
      uint64_t x = v[c * n + i] % p;
      /* subtract the contribution of earlier digits, divide by their modulus product */
      uint64_t acc = 0, prod = 1;
      for (uint64_t j = 0; j < c; j++) {
        acc = (acc + xo_mulmod(dig[j] % p, prod, p)) % p;
        prod = xo_mulmod(prod, prm->p[j] % p, p);
      }
      uint64_t diff = (x + p - acc) % p;
      dig[c] = xo_mulmod(diff, powmod(prod, p - 2, p), p);
    }
    big val, term;
    big_set(&val, 0);
    for (uint64_t c = 0; c < m; c++) {
      big_set(&term, dig[c]);
      for (uint64_t j = 0; j < c; j++) big_mul_small(&term, prm->p[j]);
      big_add(&val, &term);
    }
    big twice = val;
    big_mul_small(&twice, 2);
    u128 low = ((u128)val.l[1] << 64) | val.l[0];
    if (big_cmp(&twice, &M) > 0) low -= ((u128)M.l[1] << 64) | M.l[0]; /* value - M mod 2^128 */
    /* keep the low A bits and pack them little-endian (serializeData64/32) */
    uint64_t pos = i * A;
    for (uint32_t b = 0; b < A; b++) {
      if ((uint64_t)(low >> b) & 1) out[(pos + b) >> 3] |= (uint8_t)(1u << ((pos + b) & 7));
    }
  }
  free(v);
  return 0;
}
