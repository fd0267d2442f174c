// pir_server.cpp -- the reference's PIR server program (apps/server/main.cpp,
// apps/server/PIRServer.cpp) with the reply on the GPU.
//
// Same options and session sequence as the reference: -z/--driven [file] configures the
// parameters with a first client-driven session, keeps that session's imported database
// resident in HBM and saves the parameters, then serves every later client server-driven
// from that database; -L/--load_file file serves server-driven with saved parameters; with
// neither, every session is client-driven.  The database is the files of a directory
// (DBDirectoryProcessor, -s N to split the first file), or --db-generator.
//
// The sockets come from outside: --listen-fd N accepts clients on an inherited listening
// socket (a supervisor's, e.g. systemd socket activation), --fd N serves one client on
// an inherited connected socket (inetd-style).  Sessions are served one after another,
// or up to N at once with --concurrent N (the reference runs every session on a thread of
// its own): their uploads and downloads overlap, their GPU work shares the one GPU, and
// their database imports take turns on the database handler.
#include <errno.h>
#include <poll.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/socket.h>
#include <unistd.h>

#include <atomic>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "libpir_amd.hpp"
#include "pir_session.hpp"

using namespace xpir_amd;

namespace {

const char *kDefaultParamsFile = "exp/PIRParams.cfg";  // apps/server/main.cpp:19

void usage(const char *prog) {
  fprintf(stderr,
          "usage: %s (--listen-fd N | --fd N) [options]\n"
          "  -z, --driven [file]            configure with the first client, save to file (default %s)\n"
          "  -L, --load_file file           serve with the PIR parameters saved in file\n"
          "  -s, --split_file N             split the first file of the database directory into N elements\n"
          "  -n, --db-generator-files N     elements of the generated database (default 10)\n"
          "  -l, --db-generator-filesize B  bytes per generated element (default 12800000)\n"
          "      --db-generator             generate the database instead of reading the directory\n"
          "      --db-dir DIR               database directory (default db/)\n"
          "      --optim-dir DIR            optimizer cache directory (default exp/)\n"
          "      --no-pipeline              import the database after the query upload\n"
          "      --db-file FILE             keep the imported database's NTT words in FILE: load them from\n"
          "                                 there when FILE holds this database, else import and write it\n"
          "      --sessions K               exit after K client sessions (default: serve until killed)\n"
          "      --concurrent N             serve up to N sessions at once (default 1)\n"
          "      --device N                 GPU (default 0)\n",
          prog, kDefaultParamsFile);
}

bool number(const char *s, uint64_t *out) {
  if (!s || !*s) return false;
  char *end = nullptr;
  *out = strtoull(s, &end, 10);
  return *end == '\0';
}

}  // namespace

int main(int argc, char **argv) {
  int driven = -1;
  std::string params_file = kDefaultParamsFile, db_dir = "db/", optim_dir = "exp/", db_file;
  uint64_t split = 1, gen_n = 10, gen_l = 12800000, sessions = 0, device = 0, concurrent = 1;
  bool use_generator = false, no_pipeline = false;
  long listen_fd = -1, conn_fd = -1;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto value = [&](uint64_t *out) {
      if (i + 1 >= argc || !number(argv[i + 1], out)) {
        fprintf(stderr, "%s needs a number\n", a.c_str());
        exit(2);
      }
      i++;
    };
    auto text = [&](std::string *out) {
      if (i + 1 >= argc) {
        fprintf(stderr, "%s needs a value\n", a.c_str());
        exit(2);
      }
      *out = argv[++i];
    };
    uint64_t v = 0;
    if (a == "-h" || a == "--help") {
      usage(argv[0]);
      return 0;
    } else if (a == "-z" || a == "--driven") {
      driven = 1;
      if (i + 1 < argc && argv[i + 1][0] != '-') params_file = argv[++i];
    } else if (a == "-L" || a == "--load_file") {
      driven = 0;
      text(&params_file);
    } else if (a == "-s" || a == "--split_file") {
      value(&split);
    } else if (a == "-n" || a == "--db-generator-files") {
      value(&gen_n);
    } else if (a == "-l" || a == "--db-generator-filesize") {
      value(&gen_l);
    } else if (a == "--db-generator") {
      use_generator = true;
    } else if (a == "--db-dir") {
      text(&db_dir);
    } else if (a == "--optim-dir") {
      text(&optim_dir);
    } else if (a == "--no-pipeline") {
      no_pipeline = true;
    } else if (a == "--db-file") {
      text(&db_file);
    } else if (a == "--sessions") {
      value(&sessions);
    } else if (a == "--concurrent") {
      value(&concurrent);
      if (concurrent < 1 || concurrent > 256) {
        fprintf(stderr, "--concurrent takes 1 .. 256\n");
        return 2;
      }
    } else if (a == "--device") {
      value(&device);
    } else if (a == "--listen-fd") {
      value(&v);
      listen_fd = (long)v;
    } else if (a == "--fd") {
      value(&v);
      conn_fd = (long)v;
    } else if (a == "-p" || a == "--port") {
      fprintf(stderr, "this server takes its socket from its supervisor: use --listen-fd or --fd\n");
      return 2;
    } else {
      fprintf(stderr, "unknown option %s\n", a.c_str());
      usage(argv[0]);
      return 2;
    }
  }
  if ((listen_fd < 0) == (conn_fd < 0)) {
    usage(argv[0]);
    return 2;
  }
  if (!db_dir.empty() && db_dir.back() != '/') db_dir += '/';
  if (!optim_dir.empty() && optim_dir.back() != '/') optim_dir += '/';

  // PIRServer::PIRServer (PIRServer.cpp:26-56)
  DBHandler *db = nullptr;
  if (use_generator) {
    db = new DBGenerator(gen_n, gen_l);
  } else {
    DBDirectoryProcessor *d = split == 1 ? new DBDirectoryProcessor(db_dir) : new DBDirectoryProcessor(split, db_dir);
    if (d->getErrorStatus()) {
      fprintf(stderr, "pir_server: no database in %s\n", db_dir.c_str());
      return 1;
    }
    db = d;
  }
  fprintf(stderr, "pir_server: database of %llu elements of %llu bytes\n", (unsigned long long)db->getNbStream(),
          (unsigned long long)db->getmaxFileBytesize());
  LockedDBHandler shared_db(db);

  session_option_t opt;
  opt.driven_mode = true;
  opt.db_file = db_file;
  PIRParameters saved{};
  bool server_driven = false;
  if (driven == 0) {  // PIRServer::readPIRParamsFromFile (:116-125)
    if (readPIRParameters(saved, params_file)) {
      fprintf(stderr, "PIR parameters configuration file not found, try to launch the server in driven mode.\n");
      return 1;
    }
    server_driven = true;
    opt.driven_mode = false;
  }

  // client sessions finished (dry runs excluded); the accept loop stops at `sessions`
  std::atomic<uint64_t> served{0};
  auto done = [&]() { return sessions && served.load() >= sessions; };
  // one connected client per call: accepted, or the inherited one once; -1 when there is
  // none or the session count is reached
  bool inherited_used = false;
  auto next_client = [&]() -> int {
    if (conn_fd >= 0) {
      if (inherited_used) return -1;
      inherited_used = true;
      return (int)conn_fd;
    }
    for (;;) {
      if (done()) return -1;
      pollfd pf{(int)listen_fd, POLLIN, 0};
      const int ready = poll(&pf, 1, 100);  // wake now and then to see the count
      if (ready < 0 && errno != EINTR) return -1;
      if (ready <= 0) continue;
      const int c = accept((int)listen_fd, nullptr, nullptr);
      if (c >= 0) return c;
      if (errno != EINTR && errno != EAGAIN && errno != ECONNABORTED) return -1;
    }
  };
  // one session on a connected descriptor; what a configuration session leaves behind
  // (its parameters and kept database) is taken before the stream closes
  struct Outcome {
    bool dry_run = false, failed = false;
    PIRParameters params{};
    imported_database *kept = nullptr;
  };
  auto serve_one = [&](int fd) -> Outcome {
    FdStream stream(fd, /*owned=*/true);
    PIRSession s(stream, &shared_db, (int)device);
    s.setOptimDirectory(optim_dir);
    s.no_pipeline(no_pipeline);
    if (server_driven) s.setPIRParams(saved);
    Outcome o;
    o.dry_run = s.start(opt);
    o.failed = s.failed();
    if (o.failed) fprintf(stderr, "pir_server: session failed: %s\n", s.lastError().c_str());
    else if (!o.dry_run)
      fprintf(stderr,
              "pir_server: session done, %llu reply ciphertexts; query upload %.3f ms, import %.3f ms%s, "
              "reply %.3f ms, send %.3f ms\n",
              (unsigned long long)s.repliesSent(), s.times().upload_ms, s.times().import_ms,
              opt.got_preimported_database ? " (kept database)"
              : s.times().db_loaded        ? " (loaded from the database file)"
              : (no_pipeline ? "" : " (during the upload)"),
              s.times().reply_ms, s.times().send_ms);
    o.params = s.getPIRParams();
    o.kept = s.getSavedDatabase();
    return o;
  };

  if (driven == 1) {  // PIRServer::processDrivenSession (:70-104)
    opt.driven_mode = true;
    opt.keep_database = true;
    Outcome o;
    do {  // dry runs loop
      const int fd = next_client();
      if (fd < 0) return 1;
      delete o.kept;
      o = serve_one(fd);
    } while (o.dry_run);
    served++;
    if (o.failed) {
      fprintf(stderr, "pir_server: the configuration session failed\n");
      delete o.kept;
      return 1;
    }
    saved = o.params;
    opt.data = o.kept;
    opt.driven_mode = false;
    opt.keep_database = false;
    opt.got_preimported_database = opt.data != nullptr;
    if (writePIRParameters(saved, params_file))
      fprintf(stderr, "Unable to write PIR parameters configuration data, maybe target repertory doesn't exist ?\n");
    server_driven = true;
  }
  // PIRServer::serve / handleAccept (:58-112): a session per client, up to `concurrent`
  // at once
  std::mutex mu;
  std::condition_variable cv;
  uint64_t active = 0;
  for (;;) {
    {
      std::unique_lock<std::mutex> l(mu);
      cv.wait(l, [&] { return active < concurrent; });
    }
    const int fd = next_client();
    if (fd < 0) break;
    {
      std::lock_guard<std::mutex> l(mu);
      active++;
    }
    // detached: a long-running server must not keep a finished thread per session
    std::thread([&, fd] {
      const Outcome o = serve_one(fd);
      delete o.kept;  // only a keep_database session keeps one
      if (!o.dry_run) served++;
      std::lock_guard<std::mutex> l(mu);
      active--;
      cv.notify_all();
    }).detach();
  }
  {
    std::unique_lock<std::mutex> l(mu);
    cv.wait(l, [&] { return active == 0; });  // every session has ended
  }
  delete opt.data;
  delete db;
  return 0;
}
