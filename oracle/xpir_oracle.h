/*
 * xpir_oracle.h -- CPU restatement of XPIR's lattice (NFLLWE) PIR reply path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (xpir_amd/) never links or calls it.
 *
 * Every function restates the arithmetic of the reference (XPIR, crypto/NFLlib.*,
 * crypto/NFLLWE.cpp, pir/replyGenerator/PIRReplyGeneratorNFL_internal.cpp) from
 * its mathematics; the citation on each declaration names the reference lines
 * whose behaviour (down to the lazy [0,2p) representatives) it reproduces.
 *
 * How each part is pinned (see tests/test_oracle.py):
 *   constants/tables ... primality, root orders, table identities (closed forms)
 *   absorb ............. direct little-endian bit extraction + the reference's
 *                        own bitsplitter via oracle/_ref (straddles, A>56)
 *   NTT (lazy words) ... negacyclic DFT by brute force (residues) + the
 *                        reference's own NFLlib::ntt via oracle/_ref (exact words)
 *   MAC chain .......... sum x*y mod p (residues) + the reference's own
 *                        NFLlib::mulandaddShoup via oracle/_ref (exact words)
 *   end to end ......... decrypt(reply) == chosen element
 */
#ifndef XPIR_ORACLE_H
#define XPIR_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XO_MAX_M 4          /* 240-bit aggregated modulus at most */
#define XO_MAX_LEVELS 10    /* PIRParameters.n[MAX_REC_LVL] */

typedef struct {
  uint32_t k;        /* security bits from "LWE:k:n:bits" */
  uint32_t n;        /* polynomial degree */
  uint32_t logn;
  uint32_t m;        /* number of 60-bit moduli = bits/60 */
  uint32_t bits;     /* aggregated modulus bits */
  uint32_t Berr;     /* noise upper bound 5k/2 */
  int32_t  A;        /* absorption bits per coefficient (depends on n0) */
  uint64_t p[XO_MAX_M];
} xo_params;

typedef struct {
  uint32_t n, m;
  /* per modulus, n words each; omega tables hold n-1 used words */
  uint64_t *phi, *phiq;          /* phi^j and Shoup quotients          */
  uint64_t *w, *wq;              /* forward DIF twiddles per layer      */
  uint64_t *iw, *iwq;            /* inverse-root twiddles per layer     */
  uint64_t *ipt, *iptq;          /* n^-1 * phi^-j and quotients         */
} xo_tables;

/* P64 moduli and primitive 2^14-th roots: crypto/NFLParams.cpp:34,41 */
uint64_t xo_modulus(uint32_t i);
uint64_t xo_root(uint32_t i);
uint64_t xo_inv8192(uint32_t i);

/* "LWE:k:n:bits" -> params; A from n0 (crypto/NFLLWE.cpp:55-120, 656-668).
 * Returns 0 on success, negative on a malformed or unsupported string. */
int xo_params_init(const char *desc, uint64_t n0, xo_params *out);

/* configureNTT / prep_wtab (crypto/NFLlib.cpp:44-194, crypto/NFLlib.hpp:571-586) */
int  xo_tables_init(const xo_params *prm, xo_tables *t);
void xo_tables_free(xo_tables *t);

/* scalar primitives (crypto/NFLlib.hpp:255-321) */
uint64_t xo_mulmod(uint64_t x, uint64_t y, uint64_t p);
uint64_t xo_shoup_quot(uint64_t y, uint64_t p);
uint64_t xo_mulmod_shoup(uint64_t x, uint64_t y, uint64_t yq, uint64_t p);
uint64_t xo_mac_step(uint64_t acc, uint64_t x, uint64_t y, uint64_t yq, uint64_t p);

/* lazy Harvey DIF NTT on one n-word block (crypto/NFLlib.hpp:450-537) */
void xo_ntt(uint64_t *x, const uint64_t *w, const uint64_t *wq, uint32_t n, uint64_t p);
/* phi twist + NTT on all m blocks of one poly (crypto/NFLlib.hpp:597-608) */
void xo_ntt_powphi(const xo_tables *t, const xo_params *prm, uint64_t *poly);
/* inverse (crypto/NFLlib.hpp:539-566, 613-626) */
void xo_invntt_powinvphi(const xo_tables *t, const xo_params *prm, uint64_t *poly);

/* number of polys the absorb encoding makes of nbytes bytes (crypto/NFLlib.cpp:378) */
uint64_t xo_absorb_polys(uint64_t nbytes, uint32_t A, uint32_t n);
/* bitsplitter + CRT spread, no NTT (crypto/NFLlib.cpp:846-916); out: P*m*n words */
void xo_absorb(const uint8_t *data, uint64_t nbytes, const xo_params *prm, uint64_t *out);
/* absorb + nttAndPowPhi of every poly (crypto/NFLlib.cpp:375-401) */
void xo_import(const xo_tables *t, const xo_params *prm, const uint8_t *data, uint64_t nbytes,
               uint64_t *out);

/* One recursion level of the reply MAC (PIRReplyGeneratorNFL_internal.cpp:511-595 ->
 * crypto/NFLLWE.cpp:176-205 -> crypto/NFLlib.hpp:303-321).
 * db: (groups*n_l) elements x P x m x n lazy words; q: n_l x {a,b} x m x n canonical
 * words; out: groups x P x {a,b} x m x n.  Elements are absorbed in ascending order. */
void xo_reply_level(const xo_params *prm, const uint64_t *db, const uint64_t *q,
                    uint64_t n_l, uint64_t groups, uint32_t P, uint64_t *out, int threads);

/* xo_reply_level with the host threads split over slot blocks (all cores busy at P = 1);
 * identical words, since every chain still runs in element order on one thread.
 * mac_seconds (may be NULL) receives the time of the MAC loops alone. */
void xo_reply_level_chains(const xo_params *prm, const uint64_t *db, const uint64_t *q,
                           uint64_t n_l, uint64_t groups, uint32_t P, uint64_t *out,
                           int threads, double *mac_seconds);

/* fromResulttoInData (PIRReplyGeneratorNFL_internal.cpp:599-626): the P_l ciphertexts of
 * each group, as bytes, absorbed with the same A and transformed.  Returns P_{l+1}. */
uint32_t xo_recursion(const xo_tables *t, const xo_params *prm, const uint64_t *reply,
                      uint64_t groups, uint32_t P_l, uint64_t *out);

/* polys per element at the next level: ceil(P_l * 16nm / (A n / 8)) */
uint32_t xo_next_polys(const xo_params *prm, uint32_t P_l);

/* ---- counter-based seeded stream (SplitMix64): word k of stream `seed` ---- */
uint64_t xo_rand_word(uint64_t seed, uint64_t k);
void xo_rand_bytes(uint64_t seed, uint64_t byte_offset, uint8_t *out, uint64_t nbytes);

/* ---- test support: client side (crypto/NFLLWE.cpp:109-117, 276-372, 375-475) ---- */
/* secret key: m*n words, uniform per modulus, NTT domain */
void xo_keygen(const xo_params *prm, uint64_t key_seed, uint64_t *sk);
/* encryption of `bit` as query element number `elem` of a query with noise stream
 * `noise_seed`; out: {a,b} x m x n canonical words (wire layout) */
void xo_encrypt_bit(const xo_tables *t, const xo_params *prm, const uint64_t *sk,
                    uint32_t bit, uint64_t noise_seed, uint64_t elem, uint64_t *out);
/* decrypt one ciphertext ({a,b} x m x n) into A*n/8 plaintext bytes */
int xo_decrypt(const xo_tables *t, const xo_params *prm, const uint64_t *sk,
               const uint64_t *ct, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
