// db_directory.cpp -- host-only checks of the facade's database handlers (no GPU):
// DBDirectoryProcessor in one-file-per-element and file-splitting modes, its catalog
// strings, zero padding of short files, DBHandler::readAggregatedStream, two threads
// reading one database through LockedDBHandler, and the handlers' fingerprints.
// usage: db_directory SCRATCH_DIR   (prints "ok" and exits 0 when every check holds)
#include <fcntl.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <set>
#include <string>
#include <thread>
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
  // aggregation as DBHandler.cpp:3-26 does it: streams 0..2 in aggregates of 2.  The
  // full aggregate 0 holds streams 0 and 1; the partial aggregate 1 reads stream 2 into
  // slot 0, then zeroes one slot from slot 0 (the last real stream's), so stream 2 is
  // erased and slot 1 keeps what the buffer held
  auto elem = [&](uint64_t i) {
    const std::string &f = db.fileList()[i];
    std::vector<char> w = pattern(f == "alpha" ? 10 : f == "beta" ? 300 : 0, f == "alpha" ? 1 : f == "beta" ? 2 : 3);
    w.resize(300, 0);
    return w;
  };
  std::vector<char> agg(2 * 300, 'y');
  db.readAggregatedStream(0, 2, 0, 300, agg.data());
  CHECK(memcmp(agg.data(), elem(0).data(), 300) == 0 && memcmp(agg.data() + 300, elem(1).data(), 300) == 0);
  db.readAggregatedStream(1, 2, 0, 300, agg.data());
  for (size_t k = 0; k < 300; k++) CHECK(agg[k] == 0);
  CHECK(memcmp(agg.data() + 300, elem(1).data(), 300) == 0);
  std::vector<char> agg3(3 * 300, 'y');  // one aggregate of 3 over 3 streams: nothing zeroed
  db.readAggregatedStream(0, 3, 0, 300, agg3.data());
  for (uint64_t i = 0; i < 3; i++) CHECK(memcmp(agg3.data() + i * 300, elem(i).data(), 300) == 0);
  db.readAggregatedStream(1, 3, 0, 300, agg3.data());  // past the last stream: untouched
  CHECK(memcmp(agg3.data(), elem(0).data(), 300) == 0);

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
  // two threads importing the same many-file database through one locked handler, as
  // concurrent server sessions do: every element read equals its file, every time
  const std::string mdir = root + "many/";
  mkdir(mdir.c_str(), 0755);
  const int kFiles = 240, kBytes = 4096;
  for (int i = 0; i < kFiles; i++) {
    char name[32];
    snprintf(name, sizeof(name), "f%04d", i);
    write_file(mdir + name, pattern(kBytes, 100 + i));
  }
  DBDirectoryProcessor many(mdir);
  CHECK(!many.getErrorStatus() && many.getNbStream() == (uint64_t)kFiles);
  LockedDBHandler shared(&many);
  std::vector<std::vector<char>> want(kFiles);
  for (int i = 0; i < kFiles; i++) {
    const std::string &f = many.fileList()[i];
    want[i] = pattern(kBytes, 100 + atoi(f.c_str() + 1));
  }
  int bad[2] = {0, 0};
  auto importer = [&](int t) {
    std::vector<char> all((size_t)kFiles * kBytes);
    for (int rep = 0; rep < 20; rep++) {
      std::fill(all.begin(), all.end(), 'q');
      for (int first = 0; first < kFiles; first += 16) {
        const int cnt = std::min(16, kFiles - first);
        if (!shared.readElements(first, cnt, 0, kBytes, all.data() + (size_t)first * kBytes)) bad[t]++;
      }
      for (int i = 0; i < kFiles; i++)
        if (memcmp(all.data() + (size_t)i * kBytes, want[i].data(), kBytes) != 0) bad[t]++;
    }
  };
  std::thread t0(importer, 0), t1(importer, 1);
  t0.join();
  t1.join();
  CHECK(bad[0] == 0 && bad[1] == 0);
  CHECK(!many.readElements(kFiles, 1, 0, kBytes, agg.data()));  // no such stream: refused

  // fingerprints (what names a database file's contents): stable while nothing changes;
  // new bytes under the same name and size change it (the modification time moves)
  {
    DBDirectoryProcessor a(dir), b(dir);
    const std::string f0 = a.fingerprint();
    CHECK(f0 == b.fingerprint());
    CHECK(f0.compare(0, db.getCatalog(true).size(), db.getCatalog(true)) == 0);
    struct timespec ts[2] = {{1000000000, 0}, {1000000000, 123}};  // atime, mtime
    CHECK(utimensat(AT_FDCWD, (dir + "beta").c_str(), ts, 0) == 0);
    CHECK(a.fingerprint() != f0);
    std::vector<element_t> ev;
    std::vector<char> e0 = pattern(40, 5);
    ev.emplace_back("x", e0.size(), e0.data());
    DBVectorProcessor v(ev);
    const std::string v0 = v.fingerprint();
    e0[17] ^= 1;  // same names and sizes, other bytes
    CHECK(v.fingerprint() != v0);
    DBGenerator g1(4, 16), g2(4, 16), g3(5, 16);
    CHECK(g1.fingerprint() == g2.fingerprint() && g1.fingerprint() != g3.fingerprint());
  }

  DBDirectoryProcessor too_many(2000, sdir);  // under one byte per element
  CHECK(too_many.getErrorStatus());
  DBDirectoryProcessor missing(root + "no_such_dir/");
  CHECK(missing.getErrorStatus() && missing.getNbStream() == 0);
  if (failures) return 1;
  printf("ok\n");
  return 0;
}
