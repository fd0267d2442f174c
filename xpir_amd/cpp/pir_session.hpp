// pir_session.hpp -- the reference server's per-client protocol over any connected byte
// stream, with the reply computed on the GPU through the libpir_amd facade.
//
// Mirrors, message for message and byte for byte (native little-endian integers):
//   PIRSession::start and its steps          apps/server/PIRSession.cpp:43-100,123-565
//   PIROptimizer::controlAndCommand (server)  pir/replyGenerator/PIROptimizer.cpp:35-95,226-284
//   ServerService::read/writePIRParameters    apps/server/ServerService.cpp:24-66
//   PIRServer's session loop (-z / -L / none)  apps/server/PIRServer.cpp:26-125
// so that the reference client (apps/client/PIRClient.cpp) talks to it unchanged.  The
// socket itself is the caller's: FdStream wraps any connected descriptor (a TCP socket
// accepted by a supervisor, inetd-style, or a UNIX socket), and the server program
// (pir_server.cpp) accepts on a listening descriptor it inherits.
//
// Differences from the reference, all on error paths or in the transport: failures set
// lastError() and end the session (the reference prints and carries on with a
// half-configured session); query elements are read whole or the session fails (the
// reference warns and keeps the short element); the database import runs on a second
// thread while the query is uploaded, and the reply is sent once generated (the
// reference sends each ciphertext as its reply thread produces it).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>

#include "libpir_amd.hpp"

namespace xpir_amd {

// pir/PIRParametersExchangeMethods.hpp:21; sent as a short
enum PIRParamsExchangeMethod { CLIENT_DRIVEN, SERVER_DRIVEN };
// pir/PIROptimizerCommand.hpp:21; sent as a size_t
enum OptiCom { ABS, DATA, SPEED, EXIT, NOP };

// A connected, reliable, ordered byte stream.
class ByteStream {
 public:
  virtual ~ByteStream() {}
  // exactly `bytes` bytes, or false (end of stream or error)
  virtual bool readAll(void *buf, size_t bytes) = 0;
  virtual bool writeAll(const void *buf, size_t bytes) = 0;
  virtual void close() {}
};

// POSIX descriptors: one connected socket, or a read end and a write end.
class FdStream : public ByteStream {
 public:
  FdStream(int read_fd, int write_fd, bool owned);
  explicit FdStream(int fd, bool owned = false) : FdStream(fd, fd, owned) {}
  ~FdStream() override { close(); }
  bool readAll(void *buf, size_t bytes) override;
  bool writeAll(const void *buf, size_t bytes) override;
  // shuts a socket down for writing (the reply's end, as the reference's socket close)
  // and closes owned descriptors
  void close() override;

 private:
  int rfd_, wfd_;
  bool owned_, closed_ = false;
};

// Wall-clock milliseconds of a session's phases (no reference counterpart; the reference
// prints its PERF_TIMERS figures instead, PIRSession.cpp:384-451).
struct SessionTimes {
  double upload_ms = 0;  // first query element read .. last one read
  double import_ms = 0;  // database import (on its own thread, during the upload)
  double reply_ms = 0;   // generateReply: query to HBM, every level, reply back
  double send_ms = 0;    // reply ciphertexts written
  bool db_loaded = false; // the database came from the database file (session_option_t::db_file)
};

// apps/server/PIRSession.hpp session_option_t
struct session_option_t {
  bool driven_mode = true;                // client-driven: the client picks the parameters
  bool keep_database = false;             // keep this session's imported database
  bool got_preimported_database = false;  // `data` holds the database for these parameters
  imported_database *data = nullptr;
  // a file of the imported database's NTT words: loaded instead of importing when it holds
  // this database, else written after the import (PIRReplyGenerator::importDataCached;
  // no reference counterpart)
  std::string db_file;
};

class PIRSession {
 public:
  PIRSession(ByteStream &stream, DBHandler *db, int device = 0);
  ~PIRSession();
  // parameters a server-driven session imposes (PIRSession.cpp:594-597)
  void setPIRParams(const PIRParameters &p) { pir_ = p; }
  // the session's parameters, crypto_params as "LWE:k:n:bits:A" (PIRSession.cpp:587-591,
  // NFLLWEPublicParameters::getSerializedParams(false), NFLLWEPublicParameters.cpp:106-123)
  PIRParameters getPIRParams() const;
  // the database a keep_database session imported; ownership passes to the caller
  imported_database *getSavedDatabase();
  // directory of the optimizer's preCompute<crypto>.abs files (OptimService.cpp:20-22: "exp/")
  void setOptimDirectory(const std::string &dir) { optim_dir_ = dir; }
  // SEND_CATALOG (CMakeLists.txt:29-33, on by default): the full catalog when the
  // database has at most 1000 elements, else the compact one
  void setSendCatalog(bool full) { send_catalog_ = full; }
  // no overlap of the database import with the query upload (PIRSession.cpp:372-382)
  void no_pipeline(bool b) { no_pipeline_ = b; }
  // Runs the protocol.  true when the client was an optimizer dry run (the reference's
  // return value, which PIRServer::processDrivenSession loops on); false after a client
  // session, successful or not (see failed()).
  bool start(const session_option_t &opt);
  bool failed() const { return failed_; }
  const std::string &lastError() const { return error_; }
  uint64_t repliesSent() const { return sent_; }
  const SessionTimes &times() const { return times_; }

 private:
  bool fail(const std::string &msg);
  bool rd(void *buf, size_t bytes, const char *what);
  bool wr(const void *buf, size_t bytes, const char *what);
  bool sendCatalog();
  bool optimizerCommands();
  bool setupCrypto(const std::string &desc);
  bool sendCryptoParams();
  bool rcvCryptoParams(bool paramsandkey);
  bool rcvPirParams();
  bool sendPirParams();
  bool queryAndReply(const session_option_t &opt);

  ByteStream &s_;
  DBHandler *db_;
  int device_;
  PIRParameters pir_{};
  HomomorphicCrypto *crypto_ = nullptr;
  imported_database *saved_ = nullptr;
  std::string optim_dir_ = "exp/";
  bool send_catalog_ = true, no_pipeline_ = false, failed_ = false;
  std::string error_;
  uint64_t sent_ = 0;
  SessionTimes times_;
};

// ServerService.cpp:24-66: "d <d>\ndim <n0> <n1> ... \n<crypto_params>\n"; both return
// 0 on success and 1 on failure.  The file holds no alpha: readPIRParameters sets 1.
int writePIRParameters(const PIRParameters &p, const std::string &file_path);
int readPIRParameters(PIRParameters &p, const std::string &file_path);

}  // namespace xpir_amd
