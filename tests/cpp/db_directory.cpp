// db_directory.cpp -- host-only checks of the facade's database handlers (no GPU):
// DBDirectoryProcessor in one-file-per-element and file-splitting modes, its catalog
// strings, zero padding of short files, and DBHandler::readAggregatedStream.
// usage: db_directory SCRATCH_DIR   (prints "ok" and exits 0 when every check holds)
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <set>
#include <string>
#include <vector>

#include "../../xpir_amd/cpp/libpir_amd.hpp"

using namespace xpir_amd;

static int failures = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      failures++;                                                 \
    }                                                             \
  } while (0)

static void write_file(const std::string &path, const std::vector<char> &bytes) {
  FILE *f = fopen(path.c_str(), "wb");
  fwrite(bytes.data(), 1, bytes.size(), f);
  fclose(f);
}

static std::vector<char> pattern(size_t n, int salt) {
  std::vector<char> v(n);
  for (size_t i = 0; i < n; i++) v[i] = (char)(i * 7 + salt);
  return v;
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string root = std::string(argv[1]) + "/";
  const std::string dir = root + "db/";
  mkdir(dir.c_str(), 0755);
  write_file(dir + "alpha", pattern(10, 1));
  write_file(dir + "beta", pattern(300, 2));
  write_file(dir + "gamma", pattern(0, 3));
  write_file(dir + "-dash", pattern(5, 4));  // sorts before ".": skipped as the reference does
  mkdir((dir + "sub").c_str(), 0755);        // not a regular file: skipped

  DBDirectoryProcessor db(dir);
  CHECK(!db.getErrorStatus());
  CHECK(db.getNbStream() == 3);
  CHECK(db.getmaxFileBytesize() == 300);
  CHECK(db.getDBSizeBits() == 3 * 300 * 8);
  const std::set<std::string> names(db.fileList().begin(), db.fileList().end());
  CHECK(names == std::set<std::string>({"alpha", "beta", "gamma"}));
  // catalog strings (DBDirectoryProcessor.cpp:127-151)
  std::string full = "0\n3\n";
  for (const std::string &f : db.fileList())
    full += f + "\n" + std::to_string(f == "alpha" ? 10 : f == "beta" ? 300 : 0) + "\n";
  CHECK(db.getCatalog(true) == full);
  CHECK(db.getCatalog(false) == "1\n3\n300\n");
  // reads: data then zero padding, from an offset
  for (uint64_t i = 0; i < db.getNbStream(); i++) {
    const std::string &f = db.fileList()[i];
    const std::vector<char> want = pattern(f == "alpha" ? 10 : f == "beta" ? 300 : 0, f == "alpha" ? 1 : f == "beta" ? 2 : 3);
    std::vector<char> buf(300, 'x');
    CHECK(db.openStream(i, 4));
    CHECK(!db.openStream(i, 4));  // already open
    CHECK(db.readStream(i, buf.data(), 296) == 296);
    db.closeStream(i);
    for (size_t k = 0; k < 296; k++) CHECK(buf[k] == (k + 4 < want.size() ? want[k + 4] : 0));
    CHECK(buf[296] == 'x');
  }
  // aggregation: streams 0..2 in aggregates of 2, the missing slot zeroed
  std::vector<char> agg(2 * 300, 'y');
  db.readAggregatedStream(1, 2, 0, 300, agg.data());
  const std::string &last = db.fileList()[2];
  const std::vector<char> lw = pattern(last == "alpha" ? 10 : last == "beta" ? 300 : 0,
                                       last == "alpha" ? 1 : last == "beta" ? 2 : 3);
  for (size_t k = 0; k < 300; k++) CHECK(agg[k] == (k < lw.size() ? lw[k] : 0));
  for (size_t k = 300; k < 600; k++) CHECK(agg[k] == 0);

  // splitting: the first regular file into 7 elements of floor(1000 / 7) = 142 bytes
  const std::string sdir = root + "split/";
  mkdir(sdir.c_str(), 0755);
  const std::vector<char> big = pattern(1000, 9);
  write_file(sdir + "whole", big);
  DBDirectoryProcessor sp(7, sdir);
  CHECK(!sp.getErrorStatus());
  CHECK(sp.getNbStream() == 7 && sp.getmaxFileBytesize() == 142);
  CHECK(sp.getCatalog(false) == "1\n7\n142\n");
  CHECK(sp.getCatalog(true).rfind("0\n7\n0\n142\n1\n142\n", 0) == 0);
  std::vector<char> piece(142);
  CHECK(sp.openStream(3, 2));
  sp.readStream(3, piece.data(), 140);
  sp.closeStream(3);
  for (size_t k = 0; k < 140; k++) CHECK(piece[k] == big[3 * 142 + 2 + k]);
  // a stream read twice in a row continues where it stopped
  CHECK(sp.openStream(5, 0));
  sp.readStream(5, piece.data(), 100);
  sp.readStream(5, piece.data() + 100, 42);
  sp.closeStream(5);
  for (size_t k = 0; k < 142; k++) CHECK(piece[k] == big[5 * 142 + k]);
  // the block read of whole elements equals the element-by-element reads; other shapes decline
  std::vector<char> block(5 * 142, 'z'), one(142);
  CHECK(sp.readStreams(2, 5, 0, 142, block.data()));
  for (uint64_t i = 0; i < 5; i++) {
    CHECK(sp.openStream(2 + i, 0));
    sp.readStream(2 + i, one.data(), 142);
    sp.closeStream(2 + i);
    CHECK(memcmp(one.data(), block.data() + i * 142, 142) == 0);
  }
  CHECK(!sp.readStreams(0, 7, 1, 141, block.data()));  // from an offset: stream by stream
  CHECK(!sp.readStreams(3, 5, 0, 142, block.data()));  // past the last element
  CHECK(!db.readStreams(0, 3, 0, 300, agg.data()));    // one file per element
  DBGenerator gen(4, 16);
  std::vector<char> g(64, 0);
  CHECK(gen.readStreams(0, 4, 0, 16, g.data()));
  for (char ch : g) CHECK((unsigned char)ch == 0xaa);
  DBDirectoryProcessor too_many(2000, sdir);  // under one byte per element
  CHECK(too_many.getErrorStatus());
  DBDirectoryProcessor missing(root + "no_such_dir/");
  CHECK(missing.getErrorStatus() && missing.getNbStream() == 0);
  if (failures) return 1;
  printf("ok\n");
  return 0;
}
