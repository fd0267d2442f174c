// ref_driver.cpp -- thin C entry points over the REFERENCE's own NFLlib and database
// handlers, compiled from the unmodified sources under the reference tree by
// oracle/Makefile into oracle/_ref/libxpir_ref.so.  TEST INFRASTRUCTURE ONLY (the checker).
//
// The arithmetic all comes from the reference (crypto/NFLlib.hpp / NFLlib.cpp):
//   NFLlib::setNewParameters, deserializeDataNFL (bitsplitter + nttAndPowPhi),
//   nttAndPowPhi, invnttAndPowInvPhi, mulandaddShoup; the alpha aggregation from
//   DBHandler::readAggregatedStream over DBVectorProcessor (pir/dbhandlers/).
// The reference's NFLLWE.cpp and reply generator need Boost, which the image lacks,
// so the two loops they wrap are driven here in the reference's own order:
//   NFLLWE::mulandadd (crypto/NFLLWE.cpp:176-205): per modulus, the a chain over all
//   slots, then the b chain; and the element loop of
//   PIRReplyGeneratorNFL_internal::generateReply (PIRReplyGeneratorNFL_internal.cpp:
//   531-569): per poly (OpenMP), elements ascending.
#include "NFLlib.hpp"
#include "DBVectorProcessor.hpp"
#include <cstdint>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

struct RefHandle {
  NFLlib nfl;
  unsigned n;
  unsigned m;
};

extern "C" {

void *xr_create(unsigned n, unsigned bits) {
  RefHandle *h = new RefHandle();
  h->nfl.setNewParameters(n, bits);
  h->n = n;
  h->m = h->nfl.getnbModuli();
  return h;
}

void xr_destroy(void *hp) { delete static_cast<RefHandle *>(hp); }

// deserializeDataNFL over `nbufs` buffers of `bytes_per_buf` bytes each, stored back to
// back in `data`.  Each buffer gets its own allocation with slack, because the
// reference reads whole 64-bit words at buffer ends.  Returns the poly count, or -1 if
// `out` (cap in words) is too small.
int64_t xr_import(void *hp, unsigned A, const uint8_t *data, uint64_t nbufs,
                  uint64_t bytes_per_buf, uint64_t *out, uint64_t cap_words) {
  RefHandle *h = static_cast<RefHandle *>(hp);
  std::vector<unsigned char *> bufs(nbufs);
  for (uint64_t i = 0; i < nbufs; i++) {
    bufs[i] = static_cast<unsigned char *>(calloc(bytes_per_buf + 64, 1));
    memcpy(bufs[i], data + i * bytes_per_buf, bytes_per_buf);
  }
  uint64_t polys = 0;
  poly64 *res = h->nfl.deserializeDataNFL(bufs.data(), nbufs, bytes_per_buf * 8, A, polys);
  int64_t ret = (int64_t)polys;
  uint64_t words = polys * h->m * h->n;
  if (words > cap_words) ret = -1;
  else memcpy(out, res[0], words * sizeof(uint64_t));
  free(res[0]);
  free(res);
  for (auto b : bufs) free(b);
  return ret;
}

void xr_ntt_powphi(void *hp, uint64_t *poly) { static_cast<RefHandle *>(hp)->nfl.nttAndPowPhi(poly); }

void xr_invntt_powinvphi(void *hp, uint64_t *poly) {
  static_cast<RefHandle *>(hp)->nfl.invnttAndPowInvPhi(poly);
}

uint64_t xr_mulandadd_shoup(uint64_t rop, uint64_t x, uint64_t y, uint64_t yq, uint64_t p) {
  NFLlib::mulandaddShoup(rop, x, y, yq, p);
  return rop;
}

uint64_t xr_mulmod_shoup(uint64_t x, uint64_t y, uint64_t yq, uint64_t p) {
  return NFLlib::mulmodShoup(x, y, yq, p);
}

uint64_t xr_modulus(void *hp, unsigned i) { return static_cast<RefHandle *>(hp)->nfl.getmoduli()[i]; }

// One reply level.  db: (groups*n_l) x P x m x n; q: n_l x {a,b} x m x n (canonical);
// out: groups x P x {a,b} x m x n.  Shoup quotients as pushQuery computes them.
// mac_seconds (optional) receives the time of the MAC loops alone, the span the
// reference's own "Global reply generation" timer covers
// (PIRReplyGeneratorNFL_internal.cpp:359,444); the query Shoup precompute happens on
// pushQuery, before that span.
void xr_reply_level_timed(void *hp, const uint64_t *db, const uint64_t *q, uint64_t n_l,
                          uint64_t groups, unsigned P, uint64_t *out, int threads,
                          double *mac_seconds) {
  RefHandle *h = static_cast<RefHandle *>(hp);
  const uint64_t n = h->n, m = h->m, mn = n * m;
  const uint64_t *moduli = h->nfl.getmoduli();
  std::vector<uint64_t> qq(n_l * 2 * mn);
  for (uint64_t e = 0; e < n_l; e++)
    for (uint64_t c = 0; c < m; c++)
      for (uint64_t s = 0; s < n; s++)
        for (int ab = 0; ab < 2; ab++) {
          uint64_t idx = (e * 2 + ab) * mn + c * n + s;
          qq[idx] = ((uint128_t)q[idx] << 64) / moduli[c];
        }
  if (threads <= 0) threads = 1;
  double t0 = omp_get_wtime();
  for (uint64_t g = 0; g < groups; g++) {
#pragma omp parallel for num_threads(threads)
    for (unsigned k = 0; k < P; k++) {
      uint64_t *ra = out + (g * P + k) * 2 * mn, *rb = ra + mn;
      memset(ra, 0, 2 * mn * sizeof(uint64_t));
      for (uint64_t e = 0; e < n_l; e++) {
        const uint64_t *x = db + ((g * n_l + e) * P + k) * mn;
        const uint64_t *ya = q + e * 2 * mn, *yb = ya + mn;
        const uint64_t *yaq = qq.data() + e * 2 * mn, *ybq = yaq + mn;
        for (uint64_t c = 0; c < m; c++) {
          for (uint64_t s = 0; s < n; s++)
            NFLlib::mulandaddShoup(ra[c * n + s], x[c * n + s], ya[c * n + s], yaq[c * n + s], moduli[c]);
          for (uint64_t s = 0; s < n; s++)
            NFLlib::mulandaddShoup(rb[c * n + s], x[c * n + s], yb[c * n + s], ybq[c * n + s], moduli[c]);
        }
      }
    }
  }
  if (mac_seconds) *mac_seconds = omp_get_wtime() - t0;
}

// The alpha aggregation of importDataNFL (PIRReplyGeneratorNFL_internal.cpp:79-105): one
// rawBits buffer, calloc'd once and reused by every readAggregatedStream call in turn,
// over a DBVectorProcessor holding the elements (element i = `sizes[i]` bytes at
// data + offs[i]).  Each aggregate's rawBits content after its call is copied to
// out (nbruns x alpha * fileByteSize bytes, fileByteSize = min(bytes_per_file,
// maxFileBytesize - offset)).  Returns nbruns, or -1 if out_cap is too small.
int64_t xr_aggregate(const uint8_t *data, const uint64_t *offs, const uint64_t *sizes, uint64_t n_elems,
                     uint64_t alpha, uint64_t offset, uint64_t bytes_per_file, uint8_t *out, uint64_t out_cap) {
  std::vector<element_t> elems;
  elems.reserve(n_elems);
  for (uint64_t i = 0; i < n_elems; i++)
    elems.emplace_back(std::to_string(i), sizes[i], (char *)(data + offs[i]));
  DBVectorProcessor dbh(elems);
  const uint64_t file_bytes = std::min(bytes_per_file, dbh.getmaxFileBytesize() - offset);
  const uint64_t nbruns = (uint64_t)ceil((double)dbh.getNbStream() / alpha);
  if (nbruns * alpha * file_bytes > out_cap) return -1;
  char *raw = (char *)calloc(file_bytes * alpha, sizeof(char));
  for (uint64_t i = 0; i < nbruns; i++) {
    dbh.readAggregatedStream(i, alpha, offset, bytes_per_file, raw);
    memcpy(out + i * alpha * file_bytes, raw, alpha * file_bytes);
  }
  free(raw);
  for (auto &e : elems) free(e.data);
  return (int64_t)nbruns;
}

// Row-streamed level 0 of a d >= 2 reply over a database too large to hold: for groups
// [g0, g1), the group's n0 elements (element g*n0 + e, e ascending; elements at or past
// n_elems are the zero padding polys of importDataNFL, :146-154) are produced from the
// seeded byte stream, imported one by one with deserializeDataNFL (A bits per
// coefficient), and accumulated against the level-0 query q0 (n0 x {a,b} x m x n,
// canonical) with mulandaddShoup, each chain in element order as NFLLWE::mulandadd
// (crypto/NFLLWE.cpp:176-205) and generateReply (:531-569) run it.  out: (g1 - g0) x P0
// x {a,b} x m x n.  Groups are spread over `threads` OpenMP threads (each holds one
// group's NTT words).  Returns 0, or -1 when an element imports to other than P0 polys.
static inline uint64_t splitmix_word(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Elements [first, first + count) of the seeded database (zero polys at or past
// n_elems), each imported with deserializeDataNFL, into dst (count x P0 x m x n).
// `buf` holds elem_bytes + 64 bytes.  Returns false when an element imports to other
// than P0 polys.
static bool import_range(RefHandle *h, unsigned A, uint64_t db_seed, uint64_t n_elems, uint64_t elem_bytes,
                         uint64_t first, uint64_t count, unsigned P0, unsigned char *buf, uint64_t *dst) {
  const uint64_t mn = (uint64_t)h->n * h->m;
  bool ok = true;
  for (uint64_t e = 0; e < count; e++, dst += P0 * mn) {
    const uint64_t idx = first + e;
    if (idx >= n_elems) {
      memset(dst, 0, P0 * mn * sizeof(uint64_t));
      continue;
    }
    const uint64_t b0 = idx * elem_bytes;
    for (uint64_t j = 0; j < elem_bytes; j++) {
      const uint64_t b = b0 + j;
      buf[j] = (unsigned char)(splitmix_word(db_seed, b >> 3) >> (8 * (b & 7)));
    }
    unsigned char *bufs[1] = {buf};
    uint64_t polys = 0;
    poly64 *res = h->nfl.deserializeDataNFL(bufs, 1, elem_bytes * 8, A, polys);
    if (polys != P0)
      ok = false;
    else
      memcpy(dst, res[0], P0 * mn * sizeof(uint64_t));
    free(res[0]);
    free(res);
  }
  return ok;
}

// The NTT words of elements [first, first + count) of the seeded database, `threads`
// OpenMP threads over the elements (untimed preparation of a CPU sample).
int xr_import_range(void *hp, unsigned A, uint64_t db_seed, uint64_t n_elems, uint64_t elem_bytes, uint64_t first,
                    uint64_t count, unsigned P0, uint64_t *out, int threads) {
  RefHandle *h = static_cast<RefHandle *>(hp);
  const uint64_t mn = (uint64_t)h->n * h->m;
  const int saved = omp_get_max_active_levels();
  omp_set_max_active_levels(1);  // deserializeDataNFL's own poly loop stays serial here
  int bad = 0;
  if (threads <= 0) threads = 1;
#pragma omp parallel num_threads(threads)
  {
    unsigned char *buf = static_cast<unsigned char *>(calloc(elem_bytes + 64, 1));
#pragma omp for schedule(dynamic, 16)
    for (uint64_t e = 0; e < count; e++)
      if (!import_range(h, A, db_seed, n_elems, elem_bytes, first + e, 1, P0, buf, out + e * P0 * mn)) {
#pragma omp atomic write
        bad = 1;
      }
    free(buf);
  }
  omp_set_max_active_levels(saved);
  return bad ? -1 : 0;
}

int xr_level0_stream(void *hp, unsigned A, uint64_t db_seed, uint64_t n_elems, uint64_t elem_bytes, uint64_t n0,
                     uint64_t g0, uint64_t g1, unsigned P0, const uint64_t *q0, uint64_t *out, int threads) {
  RefHandle *h = static_cast<RefHandle *>(hp);
  const uint64_t n = h->n, m = h->m, mn = n * m;
  const uint64_t *moduli = h->nfl.getmoduli();
  std::vector<uint64_t> qq(n0 * 2 * mn);
  for (uint64_t i = 0; i < n0 * 2 * mn; i++) qq[i] = ((uint128_t)q0[i] << 64) / moduli[(i % mn) / n];
  const int saved = omp_get_max_active_levels();
  omp_set_max_active_levels(1);
  int bad = 0;
  if (threads <= 0) threads = 1;
#pragma omp parallel num_threads(threads)
  {
    std::vector<uint64_t> words(n0 * P0 * mn);
    unsigned char *buf = static_cast<unsigned char *>(calloc(elem_bytes + 64, 1));
#pragma omp for schedule(dynamic, 1)
    for (uint64_t g = g0; g < g1; g++) {
      if (!import_range(h, A, db_seed, n_elems, elem_bytes, g * n0, n0, P0, buf, words.data())) {
#pragma omp atomic write
        bad = 1;
      }
      uint64_t *og = out + (g - g0) * P0 * 2 * mn;
      memset(og, 0, P0 * 2 * mn * sizeof(uint64_t));
      for (unsigned k = 0; k < P0; k++) {
        uint64_t *ra = og + k * 2 * mn, *rb = ra + mn;
        for (uint64_t e = 0; e < n0; e++) {
          const uint64_t *x = words.data() + (e * P0 + k) * mn;
          const uint64_t *ya = q0 + e * 2 * mn, *yb = ya + mn;
          const uint64_t *yaq = qq.data() + e * 2 * mn, *ybq = yaq + mn;
          for (uint64_t c = 0; c < m; c++) {
            for (uint64_t s = 0; s < n; s++)
              NFLlib::mulandaddShoup(ra[c * n + s], x[c * n + s], ya[c * n + s], yaq[c * n + s], moduli[c]);
            for (uint64_t s = 0; s < n; s++)
              NFLlib::mulandaddShoup(rb[c * n + s], x[c * n + s], yb[c * n + s], ybq[c * n + s], moduli[c]);
          }
        }
      }
    }
    free(buf);
  }
  omp_set_max_active_levels(saved);
  return bad ? -1 : 0;
}

void xr_reply_level(void *hp, const uint64_t *db, const uint64_t *q, uint64_t n_l,
                    uint64_t groups, unsigned P, uint64_t *out, int threads) {
  xr_reply_level_timed(hp, db, q, n_l, groups, P, out, threads, nullptr);
}

}  // extern "C"
