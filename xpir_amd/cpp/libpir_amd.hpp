// libpir_amd.hpp -- C++ facade with the reference library's server-side API, running
// the reply on an MI355X through the C ABI of libxpir_amd.so.
//
// Mirrors libpir.hpp (reference root) for the NFLLWE reply path:
//   HomomorphicCryptoFactory::getCryptoMethod   libpir.hpp:52-66, pir/libpir.cpp:20-23
//   HomomorphicCrypto (the NFLLWE parameters)   crypto/HomomorphicCrypto.hpp:41-63
//   PIRParameters                               pir/PIRParameters.hpp:22-29
//   DBHandler / DBVectorProcessor / DBGenerator pir/dbhandlers/DBHandler.hpp:26-46,
//                                               DBVectorProcessor.hpp:28-60, DBGenerator.hpp
//   DBDirectoryProcessor                        pir/dbhandlers/DBDirectoryProcessor.hpp:33-66
//   imported_database                           libpir.hpp:23-33, GenericPIRReplyGenerator.hpp:34-41
//   PIRReplyGenerator                           libpir.hpp:101-169, pir/libpir.cpp:77-163
// Names, argument meaning and ownership rules are the reference's: pushQuery takes
// ownership of the query element (freed by freeQueries), popReply hands over one
// malloc'd reply element per call and returns false when none is left.  Where the
// reference exit()s or prints, this facade keeps the object usable and reports the
// status through lastStatus()/lastError().
#pragma once
#include <stdint.h>

#include <stdio.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/xpir_amd.h"

namespace xpir_amd {

#ifndef MAX_REC_LVL
#define MAX_REC_LVL 10
#endif

struct PIRParameters {
  unsigned int alpha;
  unsigned int d;
  unsigned int n[MAX_REC_LVL];
  std::string crypto_params;
};

// The NFLLWE cryptosystem's public parameters ("LWE:k:n:bits").
class HomomorphicCrypto {
 public:
  explicit HomomorphicCrypto(const std::string &desc);
  // NFLLWE::setandgetAbsBitPerCiphertext (crypto/NFLLWE.cpp:656-668)
  long setandgetAbsBitPerCiphertext(unsigned int elt_nbr);
  unsigned int getAbsorptionBitsize() const;  // A * n   (NFLLWEPublicParameters.cpp:138-141)
  unsigned int getCiphertextBitsize() const;  // 128*n*m (NFLLWEPublicParameters.cpp:148-151)
  unsigned int getpolyDegree() const { return prm_.n; }
  unsigned short getnbModuli() const { return (unsigned short)prm_.m; }
  std::string toString() const { return "LWE"; }
  const std::string &descriptor() const { return desc_; }
  const xpg_params &params() const { return prm_; }
  bool valid() const { return ok_; }
  // The client's secret key (m*n NTT-domain words), drawn on first use from a seed the
  // system's random source gives (the reference keys its generator from /dev/urandom,
  // crypto/prng/fastrandombytes.cpp:21-24), or from setKeySeed, for reproducible keys.
  void setKeySeed(uint64_t seed);
  const std::vector<uint64_t> &secretKey();

 private:
  std::string desc_;
  xpg_params prm_;
  bool ok_;
  std::vector<uint64_t> sk_;
  uint64_t key_seed_ = 0;
  bool key_seeded_ = false;
};

class HomomorphicCryptoFactory {
 public:
  // NULL when the string names no supported NFLLWE parameter set
  // (HomomorphicCryptoFactory_internal.cpp:46-65)
  static HomomorphicCrypto *getCryptoMethod(std::string cryptoParams);
};

struct Element {
  Element(std::string name_, uint64_t size_, char *data_) : name(name_), data_size(size_), data(data_) {}
  std::string name;
  size_t data_size;
  char *data;
};
typedef Element element_t;

class DBHandler {
 public:
  virtual uint64_t getNbStream() = 0;
  virtual uint64_t getmaxFileBytesize() = 0;
  virtual bool openStream(uint64_t streamNb, uint64_t requested_offset) = 0;
  virtual uint64_t readStream(uint64_t streamNb, char *buf, uint64_t size) = 0;
  virtual void closeStream(uint64_t streamNb) = 0;
  virtual ~DBHandler() {}
  // The catalog a session sends its client (PIRSession.cpp:123-152): typeOfCatalog true
  // lists the elements, "0\n<count>\n" then "<name>\n<size>\n" each; false is the compact
  // "1\n<count>\n<maxFileBytesize>\n".  This default names the elements "0", "1", ...
  // with size maxFileBytesize, as DBGenerator does (DBGenerator.cpp:29-34,46-63).
  virtual std::string getCatalog(bool typeOfCatalog);
  // Elements first .. first+count-1, `bytes` bytes each from `offset`, back to back into
  // out, when the handler reads them as one block; false: read them stream by stream.
  // (No reference counterpart: importDataNFL reads element by element, :102-135.)
  virtual bool readStreams(uint64_t first, uint64_t count, uint64_t offset, uint64_t bytes, char *out) {
    (void)first, (void)count, (void)offset, (void)bytes, (void)out;
    return false;
  }
  // Elements first .. first+count-1 into out as importData needs them: readStreams when
  // the handler reads them as one block, else stream by stream (open, read, close each).
  // False when a stream does not open.  A handler shared between threads overrides this
  // to hold its lock across the whole sequence.  (No reference counterpart.)
  virtual bool readElements(uint64_t first, uint64_t count, uint64_t offset, uint64_t bytes, char *out);
  // What identifies the database's contents for a file of its imported words
  // (PIRReplyGenerator::importDataCached): this default is the full catalog (element
  // names and sizes), which fixes the contents of generated databases; handlers whose
  // bytes can change under the same names add what changes with them.  (No reference
  // counterpart.)
  virtual std::string fingerprint() { return getCatalog(true); }
  // The alpha streams of aggregate `streamNb`, each read from `offset` and zero-padded to
  // fileByteSize = min(bytes_per_file, maxFileBytesize - offset) bytes, into slot
  // (stream mod alpha) of rawBits; when the aggregate runs past the last stream, as many
  // slots as are missing are zeroed starting at the last real stream's slot, which
  // erases that stream and leaves the slot after the zeroed ones as rawBits held it
  // (DBHandler.cpp:3-26: importDataNFL passes the same buffer for every aggregate).
  void readAggregatedStream(uint64_t streamNb, uint64_t alpha, uint64_t offset, uint64_t bytes_per_file,
                            char *rawBits);
};

// In-memory elements; short elements read as zero-padded (DBVectorProcessor.cpp:88-106).
class DBVectorProcessor : public DBHandler {
 public:
  explicit DBVectorProcessor(std::vector<element_t> &elements);
  uint64_t getNbStream() override { return elements_.size(); }
  uint64_t getmaxFileBytesize() override { return max_bytes_; }
  bool openStream(uint64_t streamNb, uint64_t requested_offset) override;
  uint64_t readStream(uint64_t streamNb, char *buf, uint64_t size) override;
  void closeStream(uint64_t streamNb) override;
  // element names and sizes (DBVectorProcessor.cpp:50-66; the compact form there lacks the
  // newline after the count, which the client's parser needs: written with it here)
  std::string getCatalog(bool typeOfCatalog) override;
  // the catalog and a hash of every element's bytes (they live in the caller's memory)
  std::string fingerprint() override;

 private:
  std::vector<element_t> &elements_;
  std::map<uint64_t, uint64_t> offsets_;
  uint64_t max_bytes_ = 0;
};

// Synthetic database of constant 0xaa bytes (DBGenerator.cpp:91-92,105).
class DBGenerator : public DBHandler {
 public:
  DBGenerator(uint64_t nbStreams, uint64_t streamBytesize, bool random = false);
  uint64_t getNbStream() override { return nb_; }
  uint64_t getmaxFileBytesize() override { return bytes_; }
  bool openStream(uint64_t, uint64_t) override { return true; }
  uint64_t readStream(uint64_t, char *buf, uint64_t size) override;
  void closeStream(uint64_t) override {}
  bool readStreams(uint64_t, uint64_t count, uint64_t, uint64_t bytes, char *out) override;

 private:
  uint64_t nb_, bytes_;
};

// Elements read from files (DBDirectoryProcessor.cpp:22-121,161-212): one element per
// file of `directory` in directory-listing order (the reference's catalog order), the
// largest file setting maxFileBytesize, shorter files read as zero-padded; or, with
// nbStreams, the first regular file split into nbStreams elements of
// floor(size / nbStreams) bytes.  As in the reference, "." and ".." and every name that
// sorts at or before "." are skipped; unlike it, only regular files are listed (the
// reference would list a subdirectory as an element).  Errors (no directory, no files,
// a file shorter than nbStreams bytes) set getErrorStatus() instead of printing.
class DBDirectoryProcessor : public DBHandler {
 public:
  explicit DBDirectoryProcessor(const std::string &directory = "db/");
  DBDirectoryProcessor(uint64_t nbStreams, const std::string &directory = "db/");
  ~DBDirectoryProcessor() override;
  // typeOfCatalog true: "0\n<count>\n" then "<name>\n<size>\n" per element; false:
  // "1\n<count>\n<maxFileBytesize>\n" (DBDirectoryProcessor.cpp:127-151)
  std::string getCatalog(bool typeOfCatalog) override;
  uint64_t getDBSizeBits() { return max_bytes_ * file_list_.size() * 8; }
  uint64_t getNbStream() override { return file_list_.size(); }
  uint64_t getmaxFileBytesize() override { return max_bytes_; }
  bool getErrorStatus() const { return error_; }
  bool openStream(uint64_t streamNb, uint64_t requested_offset) override;
  uint64_t readStream(uint64_t streamNb, char *buf, uint64_t size) override;
  void closeStream(uint64_t streamNb) override;
  // split mode, whole elements: one positional read of the file's slice
  bool readStreams(uint64_t first, uint64_t count, uint64_t offset, uint64_t bytes, char *out) override;
  static uint64_t getFileSize(const std::string &filePath);
  const std::vector<std::string> &fileList() const { return file_list_; }
  // the catalog, and each file's size and modification time (the split file's in split mode)
  std::string fingerprint() override;

 private:
  std::string directory_;
  std::vector<std::string> file_list_;
  std::map<uint64_t, FILE *> fd_pool_;
  bool splitting_ = false, error_ = false;
  std::string real_file_;  // the split file
  // split mode: one descriptor of the split file, read positionally per stream
  int split_fd_ = -1;
  std::map<uint64_t, uint64_t> split_pos_;
  uint64_t max_bytes_ = 0;
};

// A handler shared by threads (concurrent server sessions importing one database):
// every call of the inner handler under one lock, and readElements's whole
// open/read/close sequence under it at once, so that two imports never interleave on the
// per-stream state the handlers keep (DBHandler.hpp promises no thread safety;
// importDataNFL is documented as not thread-safe, PIRReplyGeneratorNFL_internal.cpp:101).
class LockedDBHandler : public DBHandler {
 public:
  explicit LockedDBHandler(DBHandler *inner) : in_(inner) {}
  uint64_t getNbStream() override { return in_->getNbStream(); }
  uint64_t getmaxFileBytesize() override { return in_->getmaxFileBytesize(); }
  bool openStream(uint64_t s, uint64_t off) override {
    std::lock_guard<std::mutex> l(mu_);
    return in_->openStream(s, off);
  }
  uint64_t readStream(uint64_t s, char *buf, uint64_t size) override {
    std::lock_guard<std::mutex> l(mu_);
    return in_->readStream(s, buf, size);
  }
  void closeStream(uint64_t s) override {
    std::lock_guard<std::mutex> l(mu_);
    in_->closeStream(s);
  }
  std::string getCatalog(bool full) override {
    std::lock_guard<std::mutex> l(mu_);
    return in_->getCatalog(full);
  }
  bool readStreams(uint64_t first, uint64_t count, uint64_t offset, uint64_t bytes, char *out) override {
    std::lock_guard<std::mutex> l(mu_);
    return in_->readStreams(first, count, offset, bytes, out);
  }
  bool readElements(uint64_t first, uint64_t count, uint64_t offset, uint64_t bytes, char *out) override {
    std::lock_guard<std::mutex> l(mu_);
    return in_->readElements(first, count, offset, bytes, out);
  }
  std::string fingerprint() override {
    std::lock_guard<std::mutex> l(mu_);
    return in_->fingerprint();
  }

 private:
  DBHandler *in_;
  std::mutex mu_;
};

// The pre-processed (NTT-domain) database, resident in HBM until destroyed.
class imported_database {
 public:
  ~imported_database();
  xpg_db *imported_database_ptr = nullptr;
  uint64_t nbElements = 0;
  uint64_t polysPerElement = 0;
  uint64_t beforeImportElementBytesize = 0;
};

// Client side, on the GPU (libpir.hpp:41-91, pir/libpir.cpp:42-69): the query for one
// aggregate of the PIR grid, as dims[l] encryptions of 0 or 1 per level
// (PIRQueryGenerator_internal.cpp:37-70), handed out one malloc'd element at a time.
class PIRQueryGenerator {
 public:
  PIRQueryGenerator(PIRParameters &pirParameters, HomomorphicCrypto &cryptoMethod, int device = 0);
  ~PIRQueryGenerator();
  // chosenElement: the aggregate's index on the grid (element index / alpha)
  void generateQuery(uint64_t chosenElement);
  // one query element (16*n*m bytes, a||b), level after level; ownership passes to the
  // caller; false when none is left
  bool popQuery(char **query);
  uint64_t getQueryElementBytesize();
  // the encryption noise stream of the next query (default: drawn from the system)
  void setNoiseSeed(uint64_t seed);
  int lastStatus() const { return status_; }
  const std::string &lastError() const { return error_; }

 private:
  int fail(int code, const std::string &msg);
  PIRParameters param_;
  HomomorphicCrypto &crypto_;
  xpg_ctx *ctx_ = nullptr;
  std::vector<char *> elems_;
  size_t next_ = 0;
  uint64_t noise_seed_ = 0;
  bool noise_seeded_ = false;
  int status_ = XPG_OK;
  std::string error_;
};

// Client side, on the GPU (libpir.hpp:171-193, pir/libpir.cpp:171-203): the reply
// ciphertexts decrypted level by level (PIRReplyExtraction_internal.cpp:41-106) into
// plaintext chunks of A*n/8 bytes, handed out one malloc'd chunk at a time.
class PIRReplyExtraction {
 public:
  PIRReplyExtraction(PIRParameters &pirParameters, HomomorphicCrypto &cryptoMethod, int device = 0);
  ~PIRReplyExtraction();
  // one reply ciphertext (16*n*m bytes) as popReply gave it; ownership passes here
  void pushEncryptedReply(char *rawBytes);
  // decrypts the pushed reply of an aggregate of alpha elements of maxFileBytesize bytes
  void extractReply(uint64_t maxFileBytesize);
  bool popPlaintextResult(char **result);
  uint64_t getPlaintextReplyBytesize();
  // ceil(maxFileBytesize * alpha / getPlaintextReplyBytesize()), in float as the reference
  uint64_t getnbPlaintextReplies(uint64_t maxFileBytesize);
  int lastStatus() const { return status_; }
  const std::string &lastError() const { return error_; }

 private:
  int fail(int code, const std::string &msg);
  PIRParameters param_;
  HomomorphicCrypto &crypto_;
  xpg_ctx *ctx_ = nullptr;
  std::vector<char *> replies_;
  std::vector<char *> chunks_;
  size_t next_ = 0;
  int status_ = XPG_OK;
  std::string error_;
};

class PIRReplyGenerator {
 public:
  PIRReplyGenerator(PIRParameters &param, HomomorphicCrypto &cryptoMethod, DBHandler *db, int device = 0);
  ~PIRReplyGenerator();
  // one query element (16*n*m bytes, a||b), level after level; ownership passes here.
  // The element goes to HBM at once (through pinned staging on a stream of its own, so
  // pushing may overlap importData on another thread) and its buffer is freed.
  void pushQuery(char *rawQuery);
  imported_database *importData(uint64_t offset, uint64_t bytes_per_db_element);
  // importData through a file of the database's NTT words (xpg_db_save / xpg_db_load; no
  // reference counterpart: the reference imports at every start).  When `path` holds
  // this database -- same handler fingerprint, offset, element bytes, alpha, grid and
  // cryptosystem -- it is loaded from there; otherwise (no file, another database or
  // parameters, a damaged file) the database is imported and the file (re)written.  The
  // file is an optimisation only: any failure to write it leaves the import in place.
  // *loaded tells which happened.  Same ownership as importData.
  imported_database *importDataCached(uint64_t offset, uint64_t bytes_per_db_element, const std::string &path,
                                      bool *loaded = nullptr);
  void generateReply(const imported_database *database);
  void freeQueries();
  bool popReply(char **reply);
  uint64_t getnbRepliesGenerated();
  uint64_t getReplyElementBytesize();
  // status of the last call (XPG_OK or an XPG_E* code) and its message
  int lastStatus() const { return status_; }
  const std::string &lastError() const { return error_; }

 private:
  int fail(int code, const std::string &msg);
  int stageFlush();
  PIRParameters param_;
  HomomorphicCrypto &crypto_;
  xpg_params prm_;
  DBHandler *db_;
  int device_;
  xpg_ctx *ctx_ = nullptr;
  bool imported_any_ = false;
  // query staging: its own context (stream), the query in HBM (every level back to back)
  // and two pinned chunks filled in turn while the other one's copy runs
  xpg_ctx *qctx_ = nullptr;
  uint64_t *qdev_ = nullptr;
  uint64_t qcap_ = 0, qbytes_ = 0, stage_fill_ = 0;
  char *stage_[2] = {nullptr, nullptr};
  int stage_cur_ = 0;
  bool stage_pending_ = false, stage_overflow_ = false;
  int stage_status_ = XPG_OK;
  std::string stage_error_;
  std::vector<char *> replies_;
  static const uint64_t kStageBytes = 32ull << 20;
  uint64_t next_reply_ = 0, generated_ = 0;
  int status_ = XPG_OK;
  std::string error_;
};

}  // namespace xpir_amd
