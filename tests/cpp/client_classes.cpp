// client_classes.cpp -- the facade's client classes on the GPU, in the manner of the
// reference's apps/simplepir/simplePIR.cpp: PIRQueryGenerator makes the query,
// PIRReplyGenerator answers it, PIRReplyExtraction decrypts the reply, and the plaintext
// chunks must spell the chosen aggregate of elements (zero-padded to the largest).
// Also: the same key and noise seeds give the same query bytes; a grid index past the
// grid and a reply of the wrong length are refused with a status.
// usage: client_classes      (prints "ok" and exits 0 when every check holds)
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

static std::vector<std::vector<char>> query_bytes(PIRParameters &p, HomomorphicCrypto &c, uint64_t chosen,
                                                  uint64_t noise) {
  PIRQueryGenerator g(p, c);
  g.setNoiseSeed(noise);
  g.generateQuery(chosen);
  std::vector<std::vector<char>> out;
  char *q = nullptr;
  while (g.popQuery(&q)) {
    out.emplace_back(q, q + g.getQueryElementBytesize());
    free(q);
  }
  return out;
}

static void run_case(const char *crypto, std::vector<unsigned> dims, unsigned alpha, uint64_t n_elems, uint64_t L,
                     uint64_t chosen) {
  fprintf(stderr, "case %s, dims %zu levels, alpha %u, %llu x %llu B, aggregate %llu\n", crypto, dims.size(), alpha,
          (unsigned long long)n_elems, (unsigned long long)L, (unsigned long long)chosen);
  std::vector<std::vector<char>> data(n_elems);
  std::vector<element_t> elements;
  for (uint64_t e = 0; e < n_elems; e++) {
    data[e].resize(L - e % 5);  // uneven: shorter elements read zero-padded
    for (size_t k = 0; k < data[e].size(); k++) data[e][k] = (char)mix(11, e * L + k);
    elements.emplace_back("e" + std::to_string(e), data[e].size(), data[e].data());
  }
  DBVectorProcessor db(elements);
  HomomorphicCrypto *c = HomomorphicCryptoFactory::getCryptoMethod(crypto);
  CHECK(c != nullptr);
  if (!c) return;
  c->setKeySeed(7);
  PIRParameters p{};
  p.alpha = alpha;
  p.d = (unsigned)dims.size();
  for (unsigned l = 0; l < p.d; l++) p.n[l] = dims[l];
  p.crypto_params = crypto;
  const uint64_t maxb = db.getmaxFileBytesize();

  PIRReplyGenerator server(p, *c, &db);
  imported_database *imported = server.importData(0, maxb);
  CHECK(imported != nullptr);
  if (!imported) {
    fprintf(stderr, "importData: %s\n", server.lastError().c_str());
    delete c;
    return;
  }
  PIRQueryGenerator qgen(p, *c);
  qgen.setNoiseSeed(3);
  qgen.generateQuery(chosen);
  CHECK(qgen.lastStatus() == XPG_OK);
  uint64_t pushed = 0, want_elems = 0;
  for (unsigned d : dims) want_elems += d;
  char *q = nullptr;
  while (qgen.popQuery(&q)) {
    server.pushQuery(q);
    pushed++;
  }
  CHECK(pushed == want_elems);
  server.generateReply(imported);
  CHECK(server.lastStatus() == XPG_OK);

  PIRReplyExtraction ext(p, *c);
  char *r = nullptr;
  while (server.popReply(&r)) ext.pushEncryptedReply(r);
  ext.extractReply(maxb);
  CHECK(ext.lastStatus() == XPG_OK);
  std::vector<char> plain;
  uint64_t chunks = 0;
  char *chunk = nullptr;
  while (ext.popPlaintextResult(&chunk)) {
    plain.insert(plain.end(), chunk, chunk + ext.getPlaintextReplyBytesize());
    free(chunk);
    chunks++;
  }
  CHECK(chunks == ext.getnbPlaintextReplies(maxb));
  CHECK(plain.size() >= maxb * alpha);
  // the aggregate as the reference's importDataNFL builds it: readAggregatedStream for
  // aggregates 0..chosen in turn over one buffer (a partial last aggregate loses its last
  // real element and keeps the previous aggregate's last slot, DBHandler.cpp:3-26)
  std::vector<char> want(maxb * alpha, 0);
  for (uint64_t u = 0; u <= chosen; u++) db.readAggregatedStream(u, alpha, 0, maxb, want.data());
  if (alpha == 3 && n_elems == 11 && chosen == 3) {  // elements 9, 10: [9, zeros, 8]
    std::vector<char> ref(maxb * alpha, 0);
    memcpy(ref.data(), data[9].data(), data[9].size());
    memcpy(ref.data() + 2 * maxb, data[8].data(), data[8].size());
    CHECK(want == ref);
  }
  CHECK(plain.size() >= want.size() && memcmp(plain.data(), want.data(), want.size()) == 0);

  // reproducible with the same seeds, different with another noise seed
  CHECK(query_bytes(p, *c, chosen, 3) == query_bytes(p, *c, chosen, 3));
  CHECK(query_bytes(p, *c, chosen, 3) != query_bytes(p, *c, chosen, 4));
  // refused: an index past the grid, a reply of the wrong length
  uint64_t units = 1;
  for (unsigned d : dims) units *= d;
  PIRQueryGenerator bad(p, *c);
  bad.generateQuery(units);
  CHECK(bad.lastStatus() == XPG_EINVAL);
  CHECK(!bad.popQuery(&q));
  PIRReplyExtraction short_reply(p, *c);
  short_reply.extractReply(maxb);  // nothing pushed
  CHECK(short_reply.lastStatus() == XPG_ESIZE);
  delete imported;
  delete c;
}

int main() {
  run_case("LWE:80:2048:120", {20}, 1, 20, 3000, 13);
  run_case("LWE:80:2048:120", {6, 5}, 1, 30, 4096, 22);
  run_case("LWE:80:1024:60", {4}, 3, 11, 2500, 3);  // the last aggregate is partial
  run_case("LWE:100:4096:180", {5, 4}, 1, 20, 2000, 17);  // three moduli, A = 78 > 64
  if (failures) return 1;
  printf("ok\n");
  return 0;
}
