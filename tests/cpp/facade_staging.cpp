// facade_staging.cpp -- the facade's query path on the GPU: PIRReplyGenerator::pushQuery
// stages every element through pinned chunks into HBM, and generateReply answers from
// there.  Each reply is compared with the C ABI's xpg_reply over the same imported
// database and the same query words (the C ABI's replies are checked against the oracle
// elsewhere).  Covers a query spanning several staging chunks, a second reply from the
// same pushed query, a new query after freeQueries, two recursion levels, and pushes of
// one element too many or too few (a size error, after which the generator still works).
// usage: facade_staging      (prints "ok" and exits 0 when every check holds)
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../xpir_amd/cpp/libpir_amd.hpp"

using namespace xpir_amd;

static int failures = 0;
#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      failures++;                                                     \
    }                                                                 \
  } while (0)

static uint64_t mix(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// canonical query words (each below its modulus), elements a||b, levels back to back
static std::vector<uint64_t> query_words(const xpg_params &p, uint64_t elems, uint64_t seed) {
  std::vector<uint64_t> q(elems * 2 * p.m * p.n);
  for (uint64_t i = 0; i < q.size(); i++) {
    const uint64_t c = (i / p.n) % p.m;
    q[i] = mix(seed, i) % p.moduli[c];
  }
  return q;
}

// pushes `elems` elements of q (malloc'd copies, ownership to the generator)
static void push(PIRReplyGenerator &g, const std::vector<uint64_t> &q, const xpg_params &p, uint64_t elems) {
  const uint64_t ct = p.ct_bytes;
  for (uint64_t e = 0; e < elems; e++) {
    char *b = static_cast<char *>(malloc(ct));
    memcpy(b, reinterpret_cast<const char *>(q.data()) + (e % (q.size() * 8 / ct)) * ct, ct);
    g.pushQuery(b);
  }
}

static std::vector<uint64_t> pop_all(PIRReplyGenerator &g) {
  std::vector<uint64_t> out;
  const uint64_t ct = g.getReplyElementBytesize();
  char *r = nullptr;
  while (g.popReply(&r)) {
    const uint64_t *w = reinterpret_cast<const uint64_t *>(r);
    out.insert(out.end(), w, w + ct / 8);
    free(r);
  }
  return out;
}

static void run_case(const char *crypto, std::vector<unsigned> dims, uint64_t n_elems, uint64_t L) {
  fprintf(stderr, "case %s, %zu levels, %llu x %llu B\n", crypto, dims.size(), (unsigned long long)n_elems,
          (unsigned long long)L);
  std::vector<std::vector<char>> data(n_elems);
  std::vector<element_t> elements;
  for (uint64_t e = 0; e < n_elems; e++) {
    data[e].resize(L);
    for (uint64_t k = 0; k < L; k++) data[e][k] = (char)mix(77, e * L + k);
    elements.emplace_back("e" + std::to_string(e), L, data[e].data());
  }
  DBVectorProcessor vec(elements);
  HomomorphicCrypto *crypto_m = HomomorphicCryptoFactory::getCryptoMethod(crypto);
  CHECK(crypto_m != nullptr);
  if (!crypto_m) return;
  PIRParameters param{};
  param.alpha = 1;
  param.d = (unsigned)dims.size();
  for (unsigned l = 0; l < param.d; l++) param.n[l] = dims[l];
  param.crypto_params = crypto;
  PIRReplyGenerator gen(param, *crypto_m, &vec);
  CHECK(gen.lastStatus() == XPG_OK);
  imported_database *db = gen.importData(0, L);
  CHECK(db != nullptr);
  if (!db) {
    fprintf(stderr, "importData: %s\n", gen.lastError().c_str());
    delete crypto_m;
    return;
  }
  const xpg_params &p = crypto_m->params();
  uint64_t elems = 0;
  for (unsigned d : dims) elems += d;
  xpg_ctx *ctx = nullptr;
  CHECK(xpg_ctx_create(0, &ctx) == XPG_OK);
  const xpg_db *h = db->imported_database_ptr;
  const uint64_t reply_words = (uint64_t)xpg_db_reply_polys(h) * 2 * p.m * p.n;
  auto expected = [&](const std::vector<uint64_t> &q) {
    std::vector<uint64_t> r(reply_words);
    CHECK(xpg_reply(ctx, h, q.data(), q.size(), r.data(), r.size()) == XPG_OK);
    return r;
  };

  const std::vector<uint64_t> q1 = query_words(p, elems, 5), q2 = query_words(p, elems, 6);
  push(gen, q1, p, elems);
  gen.generateReply(db);
  CHECK(gen.lastStatus() == XPG_OK);
  CHECK(gen.getnbRepliesGenerated() == xpg_db_reply_polys(h));
  const std::vector<uint64_t> want1 = expected(q1);
  CHECK(pop_all(gen) == want1);
  gen.generateReply(db);  // the pushed query stays until freeQueries
  CHECK(pop_all(gen) == want1);

  gen.freeQueries();
  push(gen, q2, p, elems);
  gen.generateReply(db);
  CHECK(gen.lastStatus() == XPG_OK);
  CHECK(pop_all(gen) == expected(q2));

  gen.freeQueries();
  push(gen, q1, p, elems + 1);  // one too many
  gen.generateReply(db);
  CHECK(gen.lastStatus() == XPG_ESIZE);
  gen.freeQueries();
  push(gen, q1, p, elems - 1);  // one too few
  gen.generateReply(db);
  CHECK(gen.lastStatus() == XPG_ESIZE);
  gen.freeQueries();
  push(gen, q1, p, elems);  // and the generator still answers
  gen.generateReply(db);
  CHECK(gen.lastStatus() == XPG_OK);
  CHECK(pop_all(gen) == want1);

  gen.freeQueries();
  xpg_ctx_destroy(ctx);
  delete db;
  delete crypto_m;
}

int main() {
  // 1300 elements of 64 KiB: three staging chunks of 32 MiB (512 + 512 + 276 elements)
  run_case("LWE:80:2048:120", {1300}, 1300, 200);
  // two levels, 40 + 30 elements: one chunk, a recursion step between the levels
  run_case("LWE:80:2048:120", {40, 30}, 1190, 3000);
  if (failures) return 1;
  printf("ok\n");
  return 0;
}
