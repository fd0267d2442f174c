// pir_session.cpp -- the reference server's session protocol (pir_session.hpp).
#include "pir_session.hpp"

#include <errno.h>
#include <stdlib.h>
#include <string.h>
#include <sys/socket.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <fstream>
#include <sstream>
#include <thread>
#include <vector>

namespace xpir_amd {

namespace {
double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}
const size_t kMegaByte = 1000000;                  // pir/replyGenerator/PIROptimizer.hpp:33
const uint32_t kMaxParamString = 1 << 16;          // bounds on what a peer may announce
const uint32_t kMaxKeyBytes = 1 << 24;
}  // namespace

FdStream::FdStream(int read_fd, int write_fd, bool owned) : rfd_(read_fd), wfd_(write_fd), owned_(owned) {}

bool FdStream::readAll(void *buf, size_t bytes) {
  char *p = static_cast<char *>(buf);
  while (bytes) {
    const ssize_t r = ::read(rfd_, p, bytes);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return false;
    p += r;
    bytes -= (size_t)r;
  }
  return true;
}

bool FdStream::writeAll(const void *buf, size_t bytes) {
  const char *p = static_cast<const char *>(buf);
  while (bytes) {
    // a peer that went away must not raise SIGPIPE in the server
    ssize_t r = ::send(wfd_, p, bytes, MSG_NOSIGNAL);
    if (r < 0 && errno == ENOTSOCK) r = ::write(wfd_, p, bytes);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return false;
    p += r;
    bytes -= (size_t)r;
  }
  return true;
}

void FdStream::close() {
  if (closed_) return;
  closed_ = true;
  ::shutdown(wfd_, SHUT_WR);  // ENOTSOCK for pipes: nothing to shut down
  if (owned_) {
    ::close(rfd_);
    if (wfd_ != rfd_) ::close(wfd_);
  }
}

PIRSession::PIRSession(ByteStream &stream, DBHandler *db, int device) : s_(stream), db_(db), device_(device) {
  pir_.alpha = 1;
  pir_.d = 1;
  for (unsigned i = 0; i < MAX_REC_LVL; i++) pir_.n[i] = 1;
}

PIRSession::~PIRSession() {
  delete crypto_;
  delete saved_;
}

bool PIRSession::fail(const std::string &msg) {
  if (!failed_) error_ = msg;
  failed_ = true;
  return false;
}

bool PIRSession::rd(void *buf, size_t bytes, const char *what) {
  return s_.readAll(buf, bytes) || fail(std::string("connection closed while reading ") + what);
}

bool PIRSession::wr(const void *buf, size_t bytes, const char *what) {
  return s_.writeAll(buf, bytes) || fail(std::string("connection closed while sending ") + what);
}

PIRParameters PIRSession::getPIRParams() const {
  PIRParameters p = pir_;
  if (crypto_) {
    const xpg_params &x = crypto_->params();
    p.crypto_params = "LWE:" + std::to_string(x.k) + ":" + std::to_string(x.n) + ":" + std::to_string(x.bits) +
                      ":" + std::to_string(x.A);
  }
  return p;
}

imported_database *PIRSession::getSavedDatabase() {
  imported_database *d = saved_;
  saved_ = nullptr;
  return d;
}

// PIRSession.cpp:123-152: a uint64 byte count, then the catalog text
bool PIRSession::sendCatalog() {
  const std::string buf = db_->getCatalog(send_catalog_ && db_->getNbStream() <= 1000);
  const uint64_t size = buf.size();
  return wr(&size, sizeof(size), "the catalog size") && wr(buf.data(), size, "the catalog");
}

// PIROptimizer.cpp:242-284 and the handlers above it: size_t commands until EXIT
bool PIRSession::optimizerCommands() {
  for (;;) {
    uint64_t cmd = NOP;
    if (!s_.readAll(&cmd, sizeof(cmd))) return true;  // the client left without EXIT
    switch (cmd) {
      case ABS: {
        uint32_t len = 0;
        if (!rd(&len, 4, "the cryptosystem name size")) return false;
        if (len > kMaxParamString) return fail("cryptosystem name too long");
        std::string name(len, '\0');
        if (len && !rd(&name[0], len, "the cryptosystem name")) return false;
        // OptimService::readEntireFile; a missing file goes out as an empty one, as the
        // reference's check (`< 0` on a 0/1 status) lets it
        std::string content;
        std::ifstream f(optim_dir_ + "preCompute" + name + ".abs", std::ios::binary);
        if (f.is_open()) {
          std::ostringstream ss;
          ss << f.rdbuf();
          content = ss.str();
          const size_t eof = content.find((char)EOF);
          if (eof != std::string::npos) content.resize(eof);
        }
        const int32_t size = (int32_t)content.size();
        if (!wr(&size, 4, "the optimizer cache size") || !wr(content.data(), content.size(), "the optimizer cache"))
          return false;
        break;
      }
      case DATA: {
        const uint64_t count = db_->getNbStream(), maxb = db_->getmaxFileBytesize();
        if (!wr(&count, 8, "the element count") || !wr(&maxb, 8, "the element size")) return false;
        break;
      }
      case SPEED: {
        std::vector<char> msg(kMegaByte);
        if (!rd(msg.data(), kMegaByte, "the speed test upload") || !wr(msg.data(), kMegaByte, "the speed test download"))
          return false;
        break;
      }
      case EXIT:
        return true;
      default:  // "Looping 30s as requested by client"
        std::this_thread::sleep_for(std::chrono::seconds(30));
        break;
    }
  }
}

bool PIRSession::setupCrypto(const std::string &desc) {
  // "LWE:k:n:bits" with an optional ":A" that the server recomputes from n[0]
  // (PIRReplyGeneratorNFL_internal.cpp:809)
  std::string four;
  int fields = 0;
  for (char c : desc) {
    if (c == ':' && ++fields == 4) break;
    four += c;
  }
  delete crypto_;
  crypto_ = HomomorphicCryptoFactory::getCryptoMethod(four);
  return crypto_ || fail("unsupported crypto parameters \"" + desc + "\"");
}

// PIRSession.cpp:155-172: an int byte count and the parameter string
bool PIRSession::sendCryptoParams() {
  const int32_t size = (int32_t)pir_.crypto_params.size();
  return wr(&size, 4, "the crypto parameters size") &&
         wr(pir_.crypto_params.data(), pir_.crypto_params.size(), "the crypto parameters");
}

// PIRSession.cpp:196-263: optionally the parameter string, then the key material
// (which NFLLWEPublicParameters::setModulus ignores, NFLLWEPublicParameters.cpp:44-47)
bool PIRSession::rcvCryptoParams(bool paramsandkey) {
  uint32_t size = 0;
  if (paramsandkey) {
    if (!rd(&size, 4, "the crypto parameters size")) return false;
    if (size > kMaxParamString) return fail("crypto parameter string too long");
    std::string desc(size, '\0');
    if (size && !rd(&desc[0], size, "the crypto parameters")) return false;
    if (!setupCrypto(desc)) return false;
    pir_.crypto_params = desc;
  }
  if (!rd(&size, 4, "the key size")) return false;
  if (size > kMaxKeyBytes) return fail("key material too long");
  std::vector<char> key(size);
  return size == 0 || rd(key.data(), size, "the key material");
}

// PIRSession.cpp:290-327: int d, int alpha, then d ints of dimensions
bool PIRSession::rcvPirParams() {
  int32_t d = 0, alpha = 0;
  if (!rd(&d, 4, "the recursion level") || !rd(&alpha, 4, "the aggregation")) return false;
  if (d < 1 || d > MAX_REC_LVL) return fail("recursion level out of range");
  if (alpha < 1) return fail("aggregation must be at least 1");
  pir_.d = (unsigned)d;
  pir_.alpha = (unsigned)alpha;
  for (int i = 0; i < d; i++) {
    int32_t n = 0;
    if (!rd(&n, 4, "a dimension")) return false;
    if (n < 1) return fail("dimensions must be at least 1");
    pir_.n[i] = (unsigned)n;
  }
  if (pir_.d < MAX_REC_LVL) pir_.n[pir_.d] = 1;
  return true;
}

// PIRSession.cpp:329-364
bool PIRSession::sendPirParams() {
  const int32_t d = (int32_t)pir_.d, alpha = (int32_t)pir_.alpha;
  if (!wr(&d, 4, "the recursion level") || !wr(&alpha, 4, "the aggregation")) return false;
  for (unsigned i = 0; i < pir_.d; i++) {
    const int32_t n = (int32_t)pir_.n[i];
    if (!wr(&n, 4, "a dimension")) return false;
  }
  if (pir_.d < MAX_REC_LVL) pir_.n[pir_.d] = 1;
  return true;
}

// PIRSession.cpp:370-524: query elements level after level (16*n*m bytes each, the
// query size at every recursion level, NFLLWEPublicParameters.cpp:157-160), the reply
// (from the database this session imports, or the one it was handed), then its
// ciphertexts in order and the end of the stream.
bool PIRSession::queryAndReply(const session_option_t &opt) {
  // the grid must hold every aggregate of alpha elements (the reference finds out in the
  // import; here the client learns it before uploading its query)
  const uint64_t units = (db_->getNbStream() + pir_.alpha - 1) / pir_.alpha;
  uint64_t grid = 1;
  for (unsigned l = 0; l < pir_.d; l++) grid = std::min<uint64_t>(grid * pir_.n[l], 1ull << 62);
  if (grid < units)
    return fail("the PIR grid holds " + std::to_string(grid) + " elements, the database " + std::to_string(units));
  PIRReplyGenerator gen(pir_, *crypto_, db_, device_);
  if (gen.lastStatus() != XPG_OK) return fail("reply generator: " + gen.lastError());
  const uint64_t qbytes = crypto_->params().ct_bytes;  // after setandgetAbsBitPerCiphertext(n[0])

  const bool preimported = opt.got_preimported_database && opt.data;
  imported_database *db = nullptr;
  std::string import_error;
  auto import = [&]() {
    const auto t0 = std::chrono::steady_clock::now();
    db = opt.db_file.empty() ? gen.importData(0, db_->getmaxFileBytesize())
                             : gen.importDataCached(0, db_->getmaxFileBytesize(), opt.db_file, &times_.db_loaded);
    if (!db) import_error = gen.lastError();
    times_.import_ms = ms_since(t0);
  };
  std::thread importer;
  if (!preimported && !no_pipeline_) importer = std::thread(import);

  bool ok = true;
  std::chrono::steady_clock::time_point t_up;
  for (unsigned j = 0; ok && j < pir_.d; j++)
    for (unsigned i = 0; ok && i < pir_.n[j]; i++) {
      char *buf = static_cast<char *>(malloc(qbytes));
      if (!buf) {
        ok = fail("out of host memory for the query");
        break;
      }
      ok = rd(buf, qbytes, "a query element");
      if (j == 0 && i == 0) t_up = std::chrono::steady_clock::now();  // the upload has begun
      if (ok)
        gen.pushQuery(buf);  // the generator owns it now
      else
        free(buf);
    }
  times_.upload_ms = ms_since(t_up);
  if (importer.joinable()) importer.join();
  if (!preimported && no_pipeline_ && ok) import();
  if (!ok) {
    delete db;
    return false;
  }
  if (!preimported && !db) return fail("database import: " + import_error);

  const auto t_reply = std::chrono::steady_clock::now();
  gen.generateReply(preimported ? opt.data : db);
  times_.reply_ms = ms_since(t_reply);
  if (gen.lastStatus() != XPG_OK) {
    delete db;
    return fail("reply generation: " + gen.lastError());
  }
  if (opt.keep_database && !preimported) {
    delete saved_;
    saved_ = db;
  } else {
    delete db;
  }
  const uint64_t bytes = gen.getReplyElementBytesize();
  const auto t_send = std::chrono::steady_clock::now();
  char *r = nullptr;
  while (gen.popReply(&r)) {
    const bool sent = wr(r, bytes, "a reply ciphertext");
    free(r);
    if (!sent) return false;
    sent_++;
  }
  times_.send_ms = ms_since(t_send);
  gen.freeQueries();
  return true;
}

bool PIRSession::start(const session_option_t &opt) {
  failed_ = false;
  error_.clear();
  sent_ = 0;
  times_ = SessionTimes();
  int32_t is_client = 0;
  if (!rd(&is_client, 4, "the client or optimizer choice")) {
    s_.close();
    return false;
  }
  if (is_client != 1) {  // an optimizer dry run (PIRSession.cpp:55-64)
    optimizerCommands();
    s_.close();
    return true;
  }
  const short method = opt.driven_mode ? CLIENT_DRIVEN : SERVER_DRIVEN;
  bool ok = sendCatalog() && wr(&method, sizeof(method), "the exchange method");
  if (ok && opt.driven_mode) {
    ok = optimizerCommands() && rcvCryptoParams(true) && rcvPirParams();
  } else if (ok) {
    ok = setupCrypto(pir_.crypto_params) && sendCryptoParams() && rcvCryptoParams(false) && sendPirParams();
  }
  if (ok) queryAndReply(opt);
  s_.close();
  return false;
}

int writePIRParameters(const PIRParameters &p, const std::string &file_path) {
  std::ofstream f(file_path);
  if (!f.is_open()) return 1;
  f << "d " << p.d << "\n" << "dim ";
  for (unsigned i = 0; i < p.d; i++) f << p.n[i] << " ";
  f << "\n" << p.crypto_params << "\n";
  return f.good() ? 0 : 1;
}

int readPIRParameters(PIRParameters &p, const std::string &file_path) {
  std::ifstream f(file_path);
  if (!f.is_open()) return 1;
  std::string line, tag;
  if (!std::getline(f, line)) return 1;
  std::istringstream l1(line);
  unsigned d = 0;
  if (!(l1 >> tag >> d) || tag != "d" || d < 1 || d > MAX_REC_LVL) return 1;
  if (!std::getline(f, line)) return 1;
  std::istringstream l2(line);
  if (!(l2 >> tag) || tag != "dim") return 1;
  PIRParameters q{};
  q.d = d;
  q.alpha = 1;
  for (unsigned i = 0; i < d; i++)
    if (!(l2 >> q.n[i]) || q.n[i] == 0) return 1;
  if (d < MAX_REC_LVL) q.n[d] = 1;
  if (!std::getline(f, q.crypto_params) || q.crypto_params.empty()) return 1;
  p = q;
  return 0;
}

}  // namespace xpir_amd
