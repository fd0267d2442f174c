// libpir_amd.cpp -- the C++ facade of libpir_amd.hpp over the C ABI.
#include "libpir_amd.hpp"

#include <dirent.h>
#include <errno.h>
#include <fcntl.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <random>

namespace xpir_amd {

namespace {
// Idle GPU contexts, kept for the next generator on the same device: a context keeps its
// NTT tables and grown buffers, so a session after the first starts warm.  Never freed
// (the process's end releases them); at most kPoolMax are kept.
std::mutex g_pool_mu;
std::multimap<int, xpg_ctx *> *g_pool = new std::multimap<int, xpg_ctx *>();
const size_t kPoolMax = 8;

int acquire_ctx(int device, xpg_ctx **out) {
  {
    std::lock_guard<std::mutex> l(g_pool_mu);
    auto it = g_pool->find(device);
    if (it != g_pool->end()) {
      *out = it->second;
      g_pool->erase(it);
      return XPG_OK;
    }
  }
  return xpg_ctx_create(device, out);
}

void release_ctx(int device, xpg_ctx *ctx) {
  if (!ctx) return;
  if (xpg_sync(ctx) == XPG_OK) {  // nothing of the last owner still runs on it
    std::lock_guard<std::mutex> l(g_pool_mu);
    if (g_pool->size() < kPoolMax) {
      g_pool->emplace(device, ctx);
      return;
    }
  }
  xpg_ctx_destroy(ctx);
}
}  // namespace

HomomorphicCrypto::HomomorphicCrypto(const std::string &desc) : desc_(desc) {
  memset(&prm_, 0, sizeof(prm_));
  ok_ = xpg_params_from_string(desc.c_str(), 1, &prm_) == XPG_OK;
}

long HomomorphicCrypto::setandgetAbsBitPerCiphertext(unsigned int elt_nbr) {
  xpg_params p;
  if (xpg_params_from_string(desc_.c_str(), elt_nbr, &p) != XPG_OK) return -1;
  prm_ = p;
  return prm_.A;
}

unsigned int HomomorphicCrypto::getAbsorptionBitsize() const { return (unsigned)(prm_.A * prm_.n); }
unsigned int HomomorphicCrypto::getCiphertextBitsize() const { return (unsigned)(prm_.ct_bytes * 8); }

void HomomorphicCrypto::setKeySeed(uint64_t seed) {
  key_seed_ = seed;
  key_seeded_ = true;
  sk_.clear();
}

const std::vector<uint64_t> &HomomorphicCrypto::secretKey() {
  if (sk_.empty() && ok_) {
    if (!key_seeded_) {
      std::random_device rd;
      key_seed_ = ((uint64_t)rd() << 32) ^ rd();
      key_seeded_ = true;
    }
    sk_.resize((size_t)prm_.m * prm_.n);
    if (xpg_keygen(&prm_, key_seed_, sk_.data()) != XPG_OK) sk_.clear();
  }
  return sk_;
}

// Device copy of the client's secret key on ctx (freed by the caller with xpg_dev_free).
static int upload_key(xpg_ctx *ctx, HomomorphicCrypto &crypto, uint64_t **sk_dev) {
  const std::vector<uint64_t> &sk = crypto.secretKey();
  if (sk.empty()) return XPG_EINVAL;
  void *p = nullptr;
  int rc = xpg_dev_alloc(ctx, sk.size() * 8, &p);
  if (rc == XPG_OK) rc = xpg_h2d(ctx, p, sk.data(), sk.size() * 8);
  *sk_dev = static_cast<uint64_t *>(p);
  return rc;
}

PIRQueryGenerator::PIRQueryGenerator(PIRParameters &pirParameters, HomomorphicCrypto &cryptoMethod, int device)
    : param_(pirParameters), crypto_(cryptoMethod) {
  crypto_.setandgetAbsBitPerCiphertext(param_.n[0]);
  const int rc = xpg_ctx_create(device, &ctx_);
  if (rc != XPG_OK) fail(rc, "no GPU context");
}

PIRQueryGenerator::~PIRQueryGenerator() {
  for (size_t i = next_; i < elems_.size(); i++) free(elems_[i]);
  xpg_ctx_destroy(ctx_);
}

int PIRQueryGenerator::fail(int code, const std::string &msg) {
  status_ = code;
  error_ = msg;
  if (ctx_ && code != XPG_OK) {
    const char *e = xpg_last_error(ctx_);
    if (e && *e) error_ += std::string(": ") + e;
  }
  return code;
}

void PIRQueryGenerator::setNoiseSeed(uint64_t seed) {
  noise_seed_ = seed;
  noise_seeded_ = true;
}

uint64_t PIRQueryGenerator::getQueryElementBytesize() { return crypto_.params().ct_bytes; }

void PIRQueryGenerator::generateQuery(uint64_t chosenElement) {
  status_ = XPG_OK;
  for (size_t i = next_; i < elems_.size(); i++) free(elems_[i]);
  elems_.clear();
  next_ = 0;
  if (!ctx_) {
    fail(XPG_EDEVICE, "no GPU context");
    return;
  }
  const xpg_params &prm = crypto_.params();
  uint64_t units = 1, nelems = 0;
  for (unsigned l = 0; l < param_.d; l++) {
    units *= param_.n[l];
    nelems += param_.n[l];
  }
  if (param_.d == 0 || param_.d > MAX_REC_LVL || chosenElement >= units) {
    fail(XPG_EINVAL, "chosen element outside the PIR grid");
    return;
  }
  if (!noise_seeded_) {
    std::random_device rd;
    noise_seed_ = ((uint64_t)rd() << 32) ^ rd();
  }
  noise_seeded_ = false;  // a seed set by the caller serves one query
  uint64_t *sk = nullptr;
  void *q = nullptr;
  const uint64_t bytes = nelems * prm.ct_bytes;
  std::vector<char> host(bytes);
  int rc = upload_key(ctx_, crypto_, &sk);
  if (rc == XPG_OK) rc = xpg_dev_alloc(ctx_, bytes, &q);
  if (rc == XPG_OK)
    rc = xpg_query_gen_dev(ctx_, &prm, sk, param_.d, param_.n, chosenElement, noise_seed_, static_cast<uint64_t *>(q));
  if (rc == XPG_OK) rc = xpg_d2h(ctx_, host.data(), q, bytes);
  if (q) xpg_dev_free(ctx_, q);
  if (sk) xpg_dev_free(ctx_, sk);
  if (rc != XPG_OK) {
    fail(rc, "generateQuery");
    return;
  }
  for (uint64_t e = 0; e < nelems; e++) {
    char *b = static_cast<char *>(malloc(prm.ct_bytes));
    memcpy(b, host.data() + e * prm.ct_bytes, prm.ct_bytes);
    elems_.push_back(b);
  }
}

bool PIRQueryGenerator::popQuery(char **query) {
  if (next_ >= elems_.size()) return false;
  *query = elems_[next_];
  elems_[next_++] = nullptr;
  return true;
}

PIRReplyExtraction::PIRReplyExtraction(PIRParameters &pirParameters, HomomorphicCrypto &cryptoMethod, int device)
    : param_(pirParameters), crypto_(cryptoMethod) {
  crypto_.setandgetAbsBitPerCiphertext(param_.n[0]);
  const int rc = xpg_ctx_create(device, &ctx_);
  if (rc != XPG_OK) fail(rc, "no GPU context");
}

PIRReplyExtraction::~PIRReplyExtraction() {
  for (char *r : replies_) free(r);
  for (size_t i = next_; i < chunks_.size(); i++) free(chunks_[i]);
  xpg_ctx_destroy(ctx_);
}

int PIRReplyExtraction::fail(int code, const std::string &msg) {
  status_ = code;
  error_ = msg;
  if (ctx_ && code != XPG_OK) {
    const char *e = xpg_last_error(ctx_);
    if (e && *e) error_ += std::string(": ") + e;
  }
  return code;
}

void PIRReplyExtraction::pushEncryptedReply(char *rawBytes) { replies_.push_back(rawBytes); }

uint64_t PIRReplyExtraction::getPlaintextReplyBytesize() { return crypto_.params().abs_bytes; }

uint64_t PIRReplyExtraction::getnbPlaintextReplies(uint64_t maxFileBytesize) {
  return (uint64_t)ceil((float)maxFileBytesize * param_.alpha / getPlaintextReplyBytesize());
}

void PIRReplyExtraction::extractReply(uint64_t maxFileBytesize) {
  status_ = XPG_OK;
  for (size_t i = next_; i < chunks_.size(); i++) free(chunks_[i]);
  chunks_.clear();
  next_ = 0;
  if (!ctx_) {
    fail(XPG_EDEVICE, "no GPU context");
    return;
  }
  const xpg_params &prm = crypto_.params();
  const uint64_t ct = prm.ct_bytes, ab = prm.abs_bytes;
  // ciphertexts per level, bottom up: level-l plaintexts are the level-(l-1) ciphertexts
  std::vector<uint64_t> counts{(maxFileBytesize * param_.alpha + ab - 1) / ab};
  for (unsigned l = 1; l < param_.d; l++) counts.push_back((counts.back() * ct + ab - 1) / ab);
  if (replies_.size() != counts.back()) {
    fail(XPG_ESIZE, "pushed reply ciphertexts do not match the element size");
    return;
  }
  uint64_t *sk = nullptr;
  void *cur = nullptr, *out = nullptr;
  int rc = upload_key(ctx_, crypto_, &sk);
  if (rc == XPG_OK) rc = xpg_dev_alloc(ctx_, replies_.size() * ct, &cur);
  for (size_t i = 0; rc == XPG_OK && i < replies_.size(); i++)
    rc = xpg_h2d(ctx_, static_cast<char *>(cur) + i * ct, replies_[i], ct);
  for (int l = (int)param_.d - 1; l >= 0 && rc == XPG_OK; l--) {
    rc = xpg_dev_alloc(ctx_, counts[l] * ab + 8, &out);
    if (rc == XPG_OK)
      rc = xpg_decrypt_dev(ctx_, &prm, sk, static_cast<const uint64_t *>(cur), counts[l], static_cast<uint8_t *>(out));
    xpg_dev_free(ctx_, cur);
    cur = out;
    out = nullptr;
  }
  std::vector<char> plain(counts[0] * ab);
  if (rc == XPG_OK) rc = xpg_d2h(ctx_, plain.data(), cur, plain.size());
  if (cur) xpg_dev_free(ctx_, cur);
  if (sk) xpg_dev_free(ctx_, sk);
  for (char *r : replies_) free(r);
  replies_.clear();
  if (rc != XPG_OK) {
    fail(rc, "extractReply");
    return;
  }
  const uint64_t nb = std::min<uint64_t>(getnbPlaintextReplies(maxFileBytesize), counts[0]);
  for (uint64_t i = 0; i < nb; i++) {
    char *b = static_cast<char *>(malloc(ab));
    memcpy(b, plain.data() + i * ab, ab);
    chunks_.push_back(b);
  }
}

bool PIRReplyExtraction::popPlaintextResult(char **result) {
  if (next_ >= chunks_.size()) return false;
  *result = chunks_[next_];
  chunks_[next_++] = nullptr;
  return true;
}

HomomorphicCrypto *HomomorphicCryptoFactory::getCryptoMethod(std::string cryptoParams) {
  HomomorphicCrypto *c = new HomomorphicCrypto(cryptoParams);
  if (!c->valid()) {
    delete c;
    return nullptr;
  }
  return c;
}

DBVectorProcessor::DBVectorProcessor(std::vector<element_t> &elements) : elements_(elements) {
  for (auto &e : elements_) max_bytes_ = std::max<uint64_t>(max_bytes_, e.data_size);
}

bool DBVectorProcessor::openStream(uint64_t streamNb, uint64_t requested_offset) {
  if (streamNb >= elements_.size()) return false;
  offsets_[streamNb] = std::min<uint64_t>(requested_offset, elements_[streamNb].data_size);
  return true;
}

uint64_t DBVectorProcessor::readStream(uint64_t streamNb, char *buf, uint64_t size) {
  const element_t &e = elements_[streamNb];
  uint64_t off = offsets_[streamNb];
  uint64_t avail = e.data_size - off;
  uint64_t take = std::min(avail, size);
  memcpy(buf, e.data + off, take);
  memset(buf + take, 0, size - take);
  offsets_[streamNb] = off + take;
  return size;
}

void DBVectorProcessor::closeStream(uint64_t streamNb) { offsets_.erase(streamNb); }

// Four 64-bit FNV-1a lanes with distinct offset bases over the same bytes: a 32-byte
// identity of a fingerprint string (not a cryptographic hash: it names the operator's own
// database, it does not authenticate it).
static void digest32(const std::string &text, uint8_t out[32]) {
  for (int lane = 0; lane < 4; lane++) {
    uint64_t h = 0xcbf29ce484222325ULL ^ (0x9E3779B97F4A7C15ULL * (uint64_t)(lane + 1));
    for (unsigned char c : text) h = (h ^ c) * 0x100000001b3ULL;
    for (int b = 0; b < 8; b++) out[lane * 8 + b] = (uint8_t)(h >> (8 * b));
  }
}

std::string DBVectorProcessor::fingerprint() {
  std::string fp = getCatalog(true);
  uint64_t h = 0xcbf29ce484222325ULL;
  for (const element_t &e : elements_)
    for (size_t i = 0; i < e.data_size; i++) h = (h ^ (unsigned char)e.data[i]) * 0x100000001b3ULL;
  return fp + "bytes " + std::to_string(h) + "\n";
}

std::string DBDirectoryProcessor::fingerprint() {
  std::string fp = getCatalog(true);
  auto stamp = [&](const std::string &path) {
    struct stat st;
    if (stat(path.c_str(), &st) != 0) return std::string("missing\n");
    return std::to_string((long long)st.st_size) + " " + std::to_string((long long)st.st_mtim.tv_sec) + "." +
           std::to_string((long long)st.st_mtim.tv_nsec) + "\n";
  };
  if (splitting_) return fp + real_file_ + " " + stamp(real_file_);
  for (const std::string &f : file_list_) fp += f + " " + stamp(directory_ + f);
  return fp;
}

std::string DBVectorProcessor::getCatalog(bool typeOfCatalog) {
  if (!typeOfCatalog) return "1\n" + std::to_string(getNbStream()) + "\n" + std::to_string(max_bytes_) + "\n";
  std::string buf = "0\n" + std::to_string(getNbStream()) + "\n";
  for (const element_t &e : elements_) buf += e.name + "\n" + std::to_string(e.data_size) + "\n";
  return buf;
}

std::string DBHandler::getCatalog(bool typeOfCatalog) {
  const uint64_t nb = getNbStream(), size = getmaxFileBytesize();
  if (!typeOfCatalog) return "1\n" + std::to_string(nb) + "\n" + std::to_string(size) + "\n";
  std::string buf = "0\n" + std::to_string(nb) + "\n";
  for (uint64_t i = 0; i < nb; i++) buf += std::to_string(i) + "\n" + std::to_string(size) + "\n";
  return buf;
}

void DBHandler::readAggregatedStream(uint64_t streamNb, uint64_t alpha, uint64_t offset, uint64_t bytes_per_file,
                                     char *rawBits) {
  const uint64_t maxb = getmaxFileBytesize(), nb = getNbStream();
  const uint64_t file_bytes = offset >= maxb ? 0 : std::min(bytes_per_file, maxb - offset);
  if (alpha == 0 || streamNb * alpha >= nb) return;
  const uint64_t first = streamNb * alpha, end = first + alpha;  // streams [first, end)
  const uint64_t last = std::min(end, nb) - 1;
  for (uint64_t i = first; i <= last; i++) {
    char *dst = rawBits + (i % alpha) * file_bytes;
    if (openStream(i, offset)) {
      readStream(i, dst, file_bytes);
      closeStream(i);
    } else {
      memset(dst, 0, file_bytes);
    }
  }
  const uint64_t missing = end > nb ? end - nb : 0;
  if (missing) memset(rawBits + (last % alpha) * file_bytes, 0, missing * file_bytes);
}

bool DBHandler::readElements(uint64_t first, uint64_t count, uint64_t offset, uint64_t bytes, char *out) {
  if (readStreams(first, count, offset, bytes, out)) return true;
  for (uint64_t i = first; i < first + count; i++) {
    if (!openStream(i, offset)) return false;
    readStream(i, out + (i - first) * bytes, bytes);
    closeStream(i);
  }
  return true;
}

static std::string with_slash(const std::string &d) {
  return d.empty() || d.back() == '/' ? d : d + "/";
}

static bool is_regular(const std::string &path) {
  struct stat st;
  return stat(path.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

uint64_t DBDirectoryProcessor::getFileSize(const std::string &filePath) {
  struct stat st;
  return stat(filePath.c_str(), &st) == 0 ? (uint64_t)st.st_size : 0;
}

DBDirectoryProcessor::DBDirectoryProcessor(const std::string &directory) : directory_(with_slash(directory)) {
  DIR *dir = opendir(directory_.c_str());
  if (!dir) {
    error_ = true;
    return;
  }
  while (struct dirent *ent = readdir(dir)) {
    // the reference's filter: strcmp(name, ".") > 0 && strcmp(name, "..") != 0
    if (!(strcmp(ent->d_name, ".") > 0 && strcmp(ent->d_name, "..") != 0)) continue;
    const std::string name(ent->d_name);
    if (!is_regular(directory_ + name)) continue;
    file_list_.push_back(name);
    max_bytes_ = std::max(max_bytes_, getFileSize(directory_ + name));
  }
  closedir(dir);
  if (file_list_.empty()) error_ = true;
}

DBDirectoryProcessor::DBDirectoryProcessor(uint64_t nbStreams, const std::string &directory)
    : directory_(with_slash(directory)), splitting_(true) {
  DIR *dir = opendir(directory_.c_str());
  if (!dir) {
    error_ = true;
    return;
  }
  while (struct dirent *ent = readdir(dir)) {
    const std::string name(ent->d_name);
    if (name == "." || name == ".." || !is_regular(directory_ + name)) continue;
    real_file_ = directory_ + name;  // the first regular file, as the reference splits
    break;
  }
  closedir(dir);
  const uint64_t size = real_file_.empty() ? 0 : getFileSize(real_file_);
  max_bytes_ = nbStreams ? size / nbStreams : 0;
  if (real_file_.empty() || max_bytes_ == 0) {
    error_ = true;
    return;
  }
  for (uint64_t i = 0; i < nbStreams; i++) file_list_.push_back(std::to_string(i));
}

DBDirectoryProcessor::~DBDirectoryProcessor() {
  for (auto &kv : fd_pool_) fclose(kv.second);
  if (split_fd_ >= 0) close(split_fd_);
}

std::string DBDirectoryProcessor::getCatalog(bool typeOfCatalog) {
  std::string buf;
  if (typeOfCatalog) {
    buf = "0\n" + std::to_string(getNbStream()) + "\n";
    for (const std::string &f : file_list_)
      buf += f + "\n" + std::to_string(splitting_ ? max_bytes_ : getFileSize(directory_ + f)) + "\n";
    return buf;
  }
  return "1\n" + std::to_string(getNbStream()) + "\n" + std::to_string(max_bytes_) + "\n";
}

bool DBDirectoryProcessor::openStream(uint64_t streamNb, uint64_t requested_offset) {
  if (fd_pool_.count(streamNb) || split_pos_.count(streamNb) || streamNb >= file_list_.size()) return false;
  if (splitting_) {  // every element is a slice of one file: one descriptor, no reopening
    if (split_fd_ < 0 && (split_fd_ = open(real_file_.c_str(), O_RDONLY)) < 0) return false;
    split_pos_[streamNb] = streamNb * max_bytes_ + requested_offset;
    return true;
  }
  FILE *f = fopen(splitting_ ? real_file_.c_str() : (directory_ + file_list_[streamNb]).c_str(), "rb");
  if (!f) return false;
  const uint64_t at = (splitting_ ? streamNb * max_bytes_ : 0) + requested_offset;
  if (fseeko(f, (off_t)at, SEEK_SET) != 0) {
    fclose(f);
    return false;
  }
  fd_pool_[streamNb] = f;
  return true;
}

uint64_t DBDirectoryProcessor::readStream(uint64_t streamNb, char *buf, uint64_t size) {
  uint64_t got = 0;
  auto sp = split_pos_.find(streamNb);
  if (sp != split_pos_.end()) {
    while (got < size) {
      const ssize_t r = pread(split_fd_, buf + got, size - got, (off_t)(sp->second + got));
      if (r < 0 && errno == EINTR) continue;
      if (r <= 0) break;
      got += (uint64_t)r;
    }
    sp->second += got;
  } else {
    auto it = fd_pool_.find(streamNb);
    got = it == fd_pool_.end() ? 0 : fread(buf, 1, size, it->second);
  }
  memset(buf + got, 0, size - got);  // short file: zero padding
  return size;
}

bool DBDirectoryProcessor::readStreams(uint64_t first, uint64_t count, uint64_t offset, uint64_t bytes, char *out) {
  if (!splitting_ || offset != 0 || bytes != max_bytes_ || first + count > file_list_.size()) return false;
  if (split_fd_ < 0 && (split_fd_ = open(real_file_.c_str(), O_RDONLY)) < 0) return false;
  const uint64_t total = count * bytes, at = first * max_bytes_;
  uint64_t got = 0;
  while (got < total) {
    const ssize_t r = pread(split_fd_, out + got, total - got, (off_t)(at + got));
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) break;
    got += (uint64_t)r;
  }
  memset(out + got, 0, total - got);
  return true;
}

void DBDirectoryProcessor::closeStream(uint64_t streamNb) {
  split_pos_.erase(streamNb);
  auto it = fd_pool_.find(streamNb);
  if (it == fd_pool_.end()) return;
  fclose(it->second);
  fd_pool_.erase(it);
}

DBGenerator::DBGenerator(uint64_t nbStreams, uint64_t streamBytesize, bool) : nb_(nbStreams), bytes_(streamBytesize) {}

uint64_t DBGenerator::readStream(uint64_t, char *buf, uint64_t size) {
  memset(buf, 0xaa, size);
  return size;
}

bool DBGenerator::readStreams(uint64_t, uint64_t count, uint64_t, uint64_t bytes, char *out) {
  memset(out, 0xaa, count * bytes);
  return true;
}

imported_database::~imported_database() { xpg_db_free(imported_database_ptr); }

PIRReplyGenerator::PIRReplyGenerator(PIRParameters &param, HomomorphicCrypto &cryptoMethod, DBHandler *db,
                                     int device)
    : param_(param), crypto_(cryptoMethod), db_(db), device_(device) {
  // setPirParams: A depends on n[0] (PIRReplyGeneratorNFL_internal.cpp:809-816)
  crypto_.setandgetAbsBitPerCiphertext(param_.n[0]);
  prm_ = crypto_.params();
  int rc = acquire_ctx(device, &ctx_);
  if (rc != XPG_OK) fail(rc, "no GPU context");
}

PIRReplyGenerator::~PIRReplyGenerator() {
  freeQueries();
  for (uint64_t i = next_reply_; i < replies_.size(); i++) free(replies_[i]);
  if (qctx_) {
    for (char *s : stage_)
      if (s) xpg_host_free(qctx_, s);
    if (qdev_) xpg_dev_free(qctx_, qdev_);
    release_ctx(device_, qctx_);
  }
  // a context that imported a database the caller may still hold is destroyed (which
  // detaches the database); one that only answered goes back to the pool
  if (imported_any_)
    xpg_ctx_destroy(ctx_);
  else
    release_ctx(device_, ctx_);
}

int PIRReplyGenerator::fail(int code, const std::string &msg) {
  status_ = code;
  error_ = msg;
  if (ctx_ && code != XPG_OK) {
    const char *e = xpg_last_error(ctx_);
    if (e && *e) error_ += std::string(": ") + e;
  }
  return code;
}

// Copies the filled chunk to its place in HBM without waiting and switches to the other
// chunk, whose own copy (the only one in flight, issued a chunk ago) is waited for first.
int PIRReplyGenerator::stageFlush() {
  if (stage_fill_ == 0) return XPG_OK;
  int rc;
  if (stage_pending_ && (rc = xpg_sync(qctx_)) != XPG_OK) return rc;
  const uint64_t at = qbytes_ - stage_fill_;
  rc = xpg_h2d_async(qctx_, reinterpret_cast<char *>(qdev_) + at, stage_[stage_cur_], stage_fill_);
  if (rc != XPG_OK) return rc;
  stage_pending_ = true;
  stage_cur_ ^= 1;
  stage_fill_ = 0;
  return XPG_OK;
}

void PIRReplyGenerator::pushQuery(char *rawQuery) {
  const uint64_t ct = prm_.ct_bytes;
  if (stage_status_ == XPG_OK && !qctx_) {
    uint64_t elems = 0;
    for (unsigned l = 0; l < param_.d && l < MAX_REC_LVL; l++) elems += param_.n[l];
    qcap_ = elems * ct;
    void *p = nullptr;
    if ((stage_status_ = acquire_ctx(device_, &qctx_)) != XPG_OK) {
      stage_error_ = "no GPU context for the query";
      qctx_ = nullptr;
    } else if ((stage_status_ = xpg_dev_alloc(qctx_, qcap_, &p)) != XPG_OK) {
      stage_error_ = "query allocation in HBM";
    } else {
      qdev_ = static_cast<uint64_t *>(p);
      for (char *&s : stage_) {
        void *h = nullptr;
        if ((stage_status_ = xpg_host_alloc(qctx_, kStageBytes, &h)) != XPG_OK) {
          stage_error_ = "pinned staging allocation";
          break;
        }
        s = static_cast<char *>(h);
      }
    }
  }
  if (stage_status_ == XPG_OK && qbytes_ + ct > qcap_) stage_overflow_ = true;  // more elements than the grid
  if (stage_status_ == XPG_OK && !stage_overflow_) {
    if (stage_fill_ + ct > kStageBytes && (stage_status_ = stageFlush()) != XPG_OK) stage_error_ = "query upload";
    if (stage_status_ == XPG_OK) {
      memcpy(stage_[stage_cur_] + stage_fill_, rawQuery, ct);
      stage_fill_ += ct;
      qbytes_ += ct;
    }
  }
  free(rawQuery);
}

void PIRReplyGenerator::freeQueries() {
  if (qctx_ && stage_pending_) xpg_sync(qctx_);
  qbytes_ = stage_fill_ = 0;
  stage_pending_ = stage_overflow_ = false;
}

imported_database *PIRReplyGenerator::importData(uint64_t offset, uint64_t bytes_per_db_element) {
  status_ = XPG_OK;
  if (!ctx_) {
    fail(XPG_EDEVICE, "no GPU context");
    return nullptr;
  }
  const uint64_t maxb = db_->getmaxFileBytesize();
  const uint64_t file_bytes = offset >= maxb ? 0 : std::min(bytes_per_db_element, maxb - offset);
  const uint64_t nb = db_->getNbStream();
  if (file_bytes == 0 || nb == 0) {
    fail(XPG_EINVAL, "empty database");
    return nullptr;
  }
  // elements [first, first + count) into out: one block read when the handler has one;
  // a stream that does not open fails the import (never a silently zeroed element)
  bool read_ok = true;
  auto read = [&](uint64_t first, uint64_t count, char *out) {
    if (read_ok && !db_->readElements(first, count, offset, file_bytes, out)) read_ok = false;
  };
  xpg_db *h = nullptr;
  int rc = XPG_OK;
  // the elements go to HBM through two pinned chunks, each read while the other's copy
  // runs, then the import reads them there; a plain host buffer if pinning fails
  const uint64_t per = std::max<uint64_t>(1, kStageBytes / file_bytes), chunk = per * file_bytes;
  void *pin[2] = {nullptr, nullptr}, *dev = nullptr;
  bool staged = xpg_host_alloc(ctx_, chunk, &pin[0]) == XPG_OK && xpg_host_alloc(ctx_, chunk, &pin[1]) == XPG_OK &&
                xpg_dev_alloc(ctx_, nb * file_bytes + 16, &dev) == XPG_OK;
  if (staged) {
    bool pending = false;
    int k = 0;
    for (uint64_t first = 0; first < nb && rc == XPG_OK; first += per, k ^= 1) {
      const uint64_t count = std::min(per, nb - first);
      read(first, count, static_cast<char *>(pin[k]));
      if (!read_ok) rc = XPG_EINVAL;
      if (pending) {
        const int rs = xpg_sync(ctx_);  // the other chunk's copy, issued a chunk ago
        if (rc == XPG_OK) rc = rs;
      }
      if (rc == XPG_OK)
        rc = xpg_h2d_async(ctx_, static_cast<char *>(dev) + first * file_bytes, pin[k], count * file_bytes);
      pending = true;
    }
    if (rc == XPG_OK)
      rc = xpg_db_import_dev(ctx_, &prm_, static_cast<const uint8_t *>(dev), nb, file_bytes, param_.alpha, param_.d,
                             param_.n, &h);
  }
  for (void *p : pin)
    if (p) xpg_host_free(ctx_, p);
  if (dev) xpg_dev_free(ctx_, dev);
  if (!staged) {
    std::unique_ptr<char[]> host(new (std::nothrow) char[nb * file_bytes]);  // every byte is read below
    if (!host) {
      fail(XPG_ENOMEM, "host staging of the database");
      return nullptr;
    }
    read(0, nb, host.get());
    rc = read_ok ? XPG_OK : XPG_EINVAL;
    if (rc == XPG_OK)
      rc = xpg_db_import(ctx_, &prm_, reinterpret_cast<const uint8_t *>(host.get()), nb, file_bytes, param_.alpha,
                       param_.d, param_.n, &h);
  }
  if (rc != XPG_OK) {
    fail(rc, read_ok ? "importData" : "importData: a database stream did not open");
    return nullptr;
  }
  imported_any_ = true;
  imported_database *out = new imported_database();
  out->imported_database_ptr = h;
  out->nbElements = (nb + param_.alpha - 1) / param_.alpha;
  out->polysPerElement = xpg_db_polys(h);
  out->beforeImportElementBytesize = bytes_per_db_element;
  return out;
}

imported_database *PIRReplyGenerator::importDataCached(uint64_t offset, uint64_t bytes_per_db_element,
                                                       const std::string &path, bool *loaded) {
  if (loaded) *loaded = false;
  if (!ctx_) {
    fail(XPG_EDEVICE, "no GPU context");
    return nullptr;
  }
  std::string key = db_->fingerprint() + "offset " + std::to_string(offset) + " bytes " +
                    std::to_string(bytes_per_db_element) + " alpha " + std::to_string(param_.alpha) + " d " +
                    std::to_string(param_.d) + " dims";
  for (unsigned l = 0; l < param_.d; l++) key += " " + std::to_string(param_.n[l]);
  key += " crypto " + crypto_.descriptor() + "\n";
  uint8_t digest[32];
  digest32(key, digest);
  xpg_db *h = nullptr;
  if (xpg_db_load(ctx_, &prm_, path.c_str(), digest, &h) == XPG_OK) {
    uint32_t dims[XPG_MAX_LEVELS];
    const uint32_t d = xpg_db_levels(h, dims, XPG_MAX_LEVELS);
    bool same = d == param_.d;
    for (uint32_t l = 0; same && l < d; l++) same = dims[l] == param_.n[l];
    if (same) {
      imported_any_ = true;
      imported_database *out = new imported_database();
      out->imported_database_ptr = h;
      out->nbElements = (db_->getNbStream() + param_.alpha - 1) / param_.alpha;
      out->polysPerElement = xpg_db_polys(h);
      out->beforeImportElementBytesize = bytes_per_db_element;
      if (loaded) *loaded = true;
      status_ = XPG_OK;
      return out;
    }
    xpg_db_free(h);
  }
  imported_database *out = importData(offset, bytes_per_db_element);
  if (out) (void)xpg_db_save(ctx_, out->imported_database_ptr, path.c_str(), digest);  // best effort
  return out;
}

void PIRReplyGenerator::generateReply(const imported_database *database) {
  status_ = XPG_OK;
  for (uint64_t i = next_reply_; i < replies_.size(); i++) free(replies_[i]);
  replies_.clear();
  next_reply_ = generated_ = 0;
  if (!database || !database->imported_database_ptr) {
    fail(XPG_EINVAL, "no imported database");
    return;
  }
  const xpg_db *h = database->imported_database_ptr;
  const uint64_t qwords = xpg_db_query_words(h);
  const uint64_t ct_words = 2ULL * prm_.m * prm_.n;
  if (stage_status_ != XPG_OK) {
    fail(stage_status_, stage_error_);
    return;
  }
  if (stage_overflow_ || qbytes_ != qwords * 8) {
    fail(XPG_ESIZE, "pushed query elements do not match the database grid");
    return;
  }
  int rc = stageFlush();
  if (rc == XPG_OK) rc = xpg_sync(qctx_);  // the whole query is in HBM
  if (rc != XPG_OK) {
    fail(rc, "query upload");
    return;
  }
  stage_pending_ = false;
  const uint64_t nrep = xpg_db_reply_polys(h);
  std::vector<uint64_t> r(nrep * ct_words);
  void *rdev = nullptr;
  rc = xpg_dev_alloc(ctx_, r.size() * 8, &rdev);
  if (rc == XPG_OK) rc = xpg_reply_dev(ctx_, h, qdev_, static_cast<uint64_t *>(rdev));
  if (rc == XPG_OK) rc = xpg_d2h(ctx_, r.data(), rdev, r.size() * 8);
  if (rdev) xpg_dev_free(ctx_, rdev);
  if (rc != XPG_OK) {
    fail(rc, "generateReply");
    return;
  }
  for (uint64_t i = 0; i < nrep; i++) {
    char *e = static_cast<char *>(malloc(ct_words * 8));
    memcpy(e, r.data() + i * ct_words, ct_words * 8);
    replies_.push_back(e);
  }
  generated_ = nrep;
}

bool PIRReplyGenerator::popReply(char **reply) {
  if (next_reply_ >= replies_.size()) return false;
  *reply = replies_[next_reply_];
  replies_[next_reply_++] = nullptr;
  return true;
}

uint64_t PIRReplyGenerator::getnbRepliesGenerated() { return generated_; }
uint64_t PIRReplyGenerator::getReplyElementBytesize() { return prm_.ct_bytes; }

}  // namespace xpir_amd
