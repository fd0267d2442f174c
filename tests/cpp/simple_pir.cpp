// simple_pir.cpp -- end-to-end use of the C++ facade, in the manner of the reference's
// apps/simplepir/simplePIR.cpp:3-182: factory -> PIRParameters -> DBVectorProcessor ->
// PIRReplyGenerator (pushQuery, importData, generateReply, popReply).  The query comes
// from the library's GPU query helper; the reply bytes go to a file that the test
// compares with the oracle.
//
// usage: simple_pir CRYPTO N ELEM_BYTES ALPHA D n0 [n1 ...] DB_SEED INDEX NOISE_SEED KEY_SEED OUT [DIR]
// With DIR the database is read from files through DBDirectoryProcessor: element e is
// written to DIR/e<e> with ELEM_BYTES - (e mod 3) bytes (shorter files read zero-padded),
// the catalog order is the directory listing's, and the catalog (getCatalog(true)) is
// written to OUT.catalog; INDEX is a catalog position.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include <sys/stat.h>

#include "../../xpir_amd/cpp/libpir_amd.hpp"

using namespace xpir_amd;

static uint64_t rand_word(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int main(int argc, char **argv) {
  if (argc < 12) {
    fprintf(stderr, "usage: %s CRYPTO N ELEM_BYTES ALPHA D n0.. DB_SEED INDEX NOISE_SEED KEY_SEED OUT\n", argv[0]);
    return 2;
  }
  int a = 1;
  std::string crypto = argv[a++];
  uint64_t N = strtoull(argv[a++], 0, 10), L = strtoull(argv[a++], 0, 10);
  unsigned alpha = (unsigned)atoi(argv[a++]), d = (unsigned)atoi(argv[a++]);
  PIRParameters params;
  params.alpha = alpha;
  params.d = d;
  for (unsigned i = 0; i < d; i++) params.n[i] = (unsigned)atoi(argv[a++]);
  params.crypto_params = crypto;
  uint64_t db_seed = strtoull(argv[a++], 0, 10), index = strtoull(argv[a++], 0, 10);
  uint64_t noise_seed = strtoull(argv[a++], 0, 10), key_seed = strtoull(argv[a++], 0, 10);
  const char *out_path = argv[a++];

  HomomorphicCrypto *crypto_method = HomomorphicCryptoFactory::getCryptoMethod(crypto);
  if (!crypto_method) {
    fprintf(stderr, "unsupported cryptosystem %s\n", crypto.c_str());
    return 1;
  }
  crypto_method->setandgetAbsBitPerCiphertext(params.n[0]);

  // database: N elements of L seeded bytes
  std::vector<std::vector<char>> store(N, std::vector<char>(L));
  std::vector<element_t> elements;
  for (uint64_t e = 0; e < N; e++) {
    for (uint64_t i = 0; i < L; i++) {
      uint64_t b = e * L + i;
      store[e][i] = (char)(rand_word(db_seed, b >> 3) >> (8 * (b & 7)));
    }
    elements.emplace_back("e" + std::to_string(e), L, store[e].data());
  }
  DBVectorProcessor vec(elements);
  DBHandler *db = &vec;
  DBDirectoryProcessor *dirdb = nullptr;
  if (a < argc) {
    const std::string dir = std::string(argv[a++]) + "/";
    mkdir(dir.c_str(), 0755);
    for (uint64_t e = 0; e < N; e++) {
      FILE *f = fopen((dir + "e" + std::to_string(e)).c_str(), "wb");
      if (!f) return 1;
      fwrite(store[e].data(), 1, L - (e % 3 < L ? e % 3 : 0), f);
      fclose(f);
    }
    dirdb = new DBDirectoryProcessor(dir);
    if (dirdb->getErrorStatus() || dirdb->getNbStream() != N || dirdb->getmaxFileBytesize() != L) {
      fprintf(stderr, "DBDirectoryProcessor: bad catalog\n");
      return 1;
    }
    FILE *c = fopen((std::string(out_path) + ".catalog").c_str(), "wb");
    const std::string cat = dirdb->getCatalog(true);
    fwrite(cat.data(), 1, cat.size(), c);
    fclose(c);
    db = dirdb;
  }

  PIRReplyGenerator server(params, *crypto_method, db);
  imported_database *imported = server.importData(0, L);
  if (!imported) {
    fprintf(stderr, "importData: %s\n", server.lastError().c_str());
    return 1;
  }

  // client: query elements from the GPU helper
  const xpg_params &prm = crypto_method->params();
  const uint64_t ct_words = 2ULL * prm.m * prm.n;
  uint64_t total = 0;
  for (unsigned i = 0; i < d; i++) total += params.n[i];
  xpg_ctx *cctx = nullptr;
  if (xpg_ctx_create(0, &cctx) != XPG_OK) return 1;
  std::vector<uint64_t> sk(prm.m * prm.n), query(total * ct_words);
  xpg_keygen(&prm, key_seed, sk.data());
  void *skd = nullptr, *qd = nullptr;
  xpg_dev_alloc(cctx, sk.size() * 8, &skd);
  xpg_dev_alloc(cctx, query.size() * 8, &qd);
  xpg_h2d(cctx, skd, sk.data(), sk.size() * 8);
  if (xpg_query_gen_dev(cctx, &prm, (const uint64_t *)skd, d, params.n, index, noise_seed, (uint64_t *)qd) != XPG_OK)
    return 1;
  xpg_d2h(cctx, query.data(), qd, query.size() * 8);
  xpg_dev_free(cctx, skd);
  xpg_dev_free(cctx, qd);
  xpg_ctx_destroy(cctx);
  for (uint64_t i = 0; i < total; i++) {
    char *elem = (char *)malloc(ct_words * 8);
    memcpy(elem, query.data() + i * ct_words, ct_words * 8);
    server.pushQuery(elem);  // ownership passes to the generator
  }

  server.generateReply(imported);
  if (server.lastStatus() != XPG_OK) {
    fprintf(stderr, "generateReply: %s\n", server.lastError().c_str());
    return 1;
  }
  FILE *f = fopen(out_path, "wb");
  char *reply = nullptr;
  uint64_t count = 0;
  while (server.popReply(&reply)) {
    fwrite(reply, 1, server.getReplyElementBytesize(), f);
    free(reply);
    count++;
  }
  fclose(f);
  server.freeQueries();
  delete imported;
  delete crypto_method;
  delete dirdb;
  printf("simple_pir: %llu reply elements of %llu bytes\n", (unsigned long long)count,
         (unsigned long long)server.getReplyElementBytesize());
  return count == server.getnbRepliesGenerated() ? 0 : 1;
}
