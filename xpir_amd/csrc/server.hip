// server.hip -- continuous-batching reply server over resident databases.
//
// The reference serves one client per session thread and one query per generator
// (apps/server/PIRSession.cpp:371-493).  Here one dispatcher thread owns W worker
// contexts (one HIP stream each).  The server's device memory is one pool of
// reply_pool_bytes: half of it holds reply slots, the other half is split into one
// workspace per worker, from which a batch's intermediate levels (level-0 outputs, level
// databases, block summaries) are carved.  A request is admitted when its reply slot fits
// (else XPG_EBUSY: try again after releasing completed replies); one whose intermediates
// alone exceed a worker's workspace can never run (XPG_ESIZE).  Pending requests are
// grouped per database, as many as the worker's workspace holds (at most max_batch),
// and launched as one batched reply on the first idle worker, so arrivals during a
// running batch are picked up by the next free worker and their kernels overlap on the
// GPU.  Completion is detected with HIP events; latency is measured from submit to the
// event that ends the request's batch.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <set>
#include <thread>

#include "internal.hpp"

namespace {

struct Req {
  uint64_t ticket;
  const xpg_db *db;
  const uint64_t *query;
  uint64_t *reply;
  uint64_t reply_words, footprint;
  double submit_us, start_us;
};

struct Worker {
  xpg_ctx *ctx = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  bool heavy = false;     // its batch has more exact steps than the heavy threshold
  std::vector<Req> batch;
  uint8_t *ws = nullptr;  // this worker's share of the pool
};

// first-fit allocator over one device slab (host bookkeeping only)
struct Slab {
  uint8_t *base = nullptr;
  uint64_t size = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> length
  bool alloc(uint64_t len, uint64_t *off) {
    len = (len + 255) & ~255ULL;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= len) {
        *off = it->first;
        const uint64_t rest = it->second - len, o = it->first + len;
        free_.erase(it);
        if (rest) free_[o] = rest;
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t len) {
    len = (len + 255) & ~255ULL;
    auto it = free_.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

double now_us() {
  using namespace std::chrono;
  return duration<double, std::micro>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct xpg_server {
  int device = 0;
  xpg_server_config cfg{};
  std::vector<Worker> workers;
  uint8_t *pool = nullptr;  // reply_pool_bytes: the reply slots, then the workspaces
  uint64_t ws_bytes = 0;    // each worker's workspace
  uint64_t heavy_steps = 0; // heavy batch threshold (exact chain steps)
  Slab slab;                // the reply slots
  uint64_t budget_used = 0;
  uint64_t next_ticket = 1;
  std::deque<Req> pending;
  std::vector<xpg_completion> done;
  std::map<uint64_t, std::pair<uint64_t, uint64_t>> slots;  // ticket -> (offset, bytes) of its reply
  std::map<uint64_t, uint64_t> footprint;                   // ticket -> admitted bytes
  std::set<uint64_t> finished;                              // tickets whose completion is posted
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::thread th;
  bool stop = false;
  uint64_t batches = 0, batched_requests = 0;
};

// exact chain steps of one request's reply, every level
static uint64_t request_steps(const xpg_db *db) {
  const uint64_t mn2 = 2ull * db->prm.m * db->prm.n;
  uint64_t s = 0;
  for (uint32_t l = 0; l < db->d; l++) {
    uint64_t groups = 1;
    for (uint32_t j = l + 1; j < db->d; j++) groups *= db->dims[j];
    s += groups * db->Pl[l] * mn2 * db->dims[l];
  }
  return s;
}

// a request's reply slot (its intermediates come from the worker's workspace)
static uint64_t request_footprint(const xpg_db *db, uint64_t *reply_words) {
  const uint64_t mn = (uint64_t)db->prm.m * db->prm.n;
  *reply_words = (uint64_t)db->Pl[db->d] * 2 * mn;
  return (*reply_words * 8 + 255) & ~255ULL;
}

static void post_completion(xpg_server *s, const Req &r, double t, int status) {
  xpg_completion c{};
  c.ticket = r.ticket;
  c.reply_dev = r.reply;
  c.reply_words = r.reply_words;
  c.submit_us = r.submit_us;
  c.start_us = r.start_us;
  c.done_us = t;
  c.status = status;
  s->done.push_back(c);
  s->finished.insert(r.ticket);
}

static void dispatcher(xpg_server *s) {
  std::unique_lock<std::mutex> lk(s->mu);
  (void)hipSetDevice(s->device);
  while (!s->stop) {
    bool progressed = false;
    // completions
    for (Worker &w : s->workers) {
      if (!w.busy || hipEventQuery(w.done) != hipSuccess) continue;
      const double t = now_us();
      for (const Req &r : w.batch) post_completion(s, r, t, XPG_OK);
      w.batch.clear();
      w.busy = false;
      progressed = true;
    }
    if (progressed) s->done_cv.notify_all();
    // dispatch: the database of the oldest pending request, up to max_batch of its
    // requests and as many as the worker's workspace holds -- unless that batch is heavy
    // while another worker runs a heavy one: then the oldest light batch, or nothing
    for (Worker &w : s->workers) {
      if (w.busy || s->pending.empty()) continue;
      bool heavy_running = false;
      for (const Worker &o : s->workers) heavy_running |= o.busy && o.heavy;
      const xpg_db *db = nullptr;
      uint32_t fit = 0;
      bool heavy = false;
      std::map<const xpg_db *, uint32_t> waiting;  // pending requests per database
      for (const Req &q : s->pending) waiting[q.db]++;
      std::set<const xpg_db *> seen;
      for (const Req &r : s->pending) {
        if (seen.size() == waiting.size()) break;  // every database considered
        if (!seen.insert(r.db).second) continue;
        const uint32_t wait_n = waiting[r.db];
        uint32_t f = 1;
        while (f < s->cfg.max_batch && f < wait_n && batch_workspace_bytes(w.ctx, r.db, f + 1) <= s->ws_bytes) f++;
        const bool h = request_steps(r.db) * f > s->heavy_steps;
        if (h && heavy_running) continue;
        db = r.db;
        fit = f;
        heavy = h;
        break;
      }
      if (!db) continue;
      std::vector<Req> take;
      for (auto it = s->pending.begin(); it != s->pending.end() && take.size() < fit;) {
        if (it->db == db) {
          take.push_back(*it);
          it = s->pending.erase(it);
        } else {
          ++it;
        }
      }
      const double t = now_us();
      std::vector<xpg_request> reqs(take.size());
      for (size_t i = 0; i < take.size(); i++) {
        take[i].start_us = t;
        reqs[i] = xpg_request{take[i].db, take[i].query, take[i].reply};
      }
      lk.unlock();
      int rc = reply_batch_in_workspace(w.ctx, reqs.data(), (uint32_t)reqs.size(), w.ws, s->ws_bytes);
      if (rc == XPG_OK && hipEventRecord(w.done, w.ctx->stream) != hipSuccess) rc = XPG_EDEVICE;
      // a failed batch may have launched kernels already: they finish before its slots
      // and workspace are handed back
      if (rc != XPG_OK) (void)hipStreamSynchronize(w.ctx->stream);
      lk.lock();
      s->batches++;
      s->batched_requests += take.size();
      if (rc != XPG_OK) {
        const double t1 = now_us();
        for (const Req &r : take) post_completion(s, r, t1, rc);
        s->done_cv.notify_all();
      } else {
        w.batch = std::move(take);
        w.busy = true;
        w.heavy = heavy;
      }
      progressed = true;
    }
    if (!progressed) {
      bool any_busy = false;
      for (Worker &w : s->workers) any_busy |= w.busy;
      // busy workers: poll their events; idle: sleep until a submit arrives
      if (any_busy) s->cv.wait_for(lk, std::chrono::microseconds(20));
      else s->cv.wait_for(lk, std::chrono::milliseconds(5));
    }
  }
}

extern "C" {

int xpg_server_create(int device, const xpg_server_config *cfg, xpg_server **out) {
  if (!cfg || !out || cfg->workers == 0 || cfg->max_batch == 0 || cfg->reply_pool_bytes == 0) return XPG_EINVAL;
  *out = nullptr;
  xpg_server *s = new xpg_server();
  s->device = device;
  s->cfg = *cfg;
  s->heavy_steps = cfg->heavy_steps ? cfg->heavy_steps : 1000000000ull;
  for (uint32_t i = 0; i < cfg->workers; i++) {
    Worker w;
    if (xpg_ctx_create(device, &w.ctx) != XPG_OK || hipEventCreateWithFlags(&w.done, hipEventDisableTiming) != hipSuccess) {
      if (w.ctx) xpg_ctx_destroy(w.ctx);
      for (Worker &x : s->workers) {
        (void)hipEventDestroy(x.done);
        xpg_ctx_destroy(x.ctx);
      }
      delete s;
      return XPG_EDEVICE;
    }
    s->workers.push_back(w);
  }
  (void)hipSetDevice(device);
  // half the pool for reply slots, the other half one workspace per worker
  s->ws_bytes = (cfg->reply_pool_bytes / 2 / cfg->workers) & ~255ULL;
  const uint64_t slots = (cfg->reply_pool_bytes - s->ws_bytes * cfg->workers) & ~255ULL;
  const int prc = s->ws_bytes == 0 || slots == 0 ? XPG_EINVAL
                  : hipMalloc((void **)&s->pool, cfg->reply_pool_bytes) != hipSuccess ? XPG_ENOMEM
                                                                                      : XPG_OK;
  if (prc != XPG_OK) {
    for (Worker &x : s->workers) {
      (void)hipEventDestroy(x.done);
      xpg_ctx_destroy(x.ctx);
    }
    delete s;
    return prc;
  }
  s->slab.base = s->pool;
  s->slab.size = slots;
  s->slab.free_[0] = slots;
  for (uint32_t i = 0; i < cfg->workers; i++) s->workers[i].ws = s->pool + slots + i * s->ws_bytes;
  s->th = std::thread(dispatcher, s);
  *out = s;
  return XPG_OK;
}

void xpg_server_destroy(xpg_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  s->th.join();
  for (Worker &w : s->workers) {
    (void)hipStreamSynchronize(w.ctx->stream);
    (void)hipEventDestroy(w.done);
    xpg_ctx_destroy(w.ctx);
  }
  (void)hipFree(s->pool);
  delete s;
}

int xpg_server_submit(xpg_server *s, const xpg_db *db, const uint64_t *query_dev, uint64_t *ticket) {
  if (!s || !db || !query_dev || !ticket) return XPG_EINVAL;
  if (db->device != s->device) return XPG_EINVAL;
  Req r{};
  r.db = db;
  r.query = query_dev;
  r.footprint = request_footprint(db, &r.reply_words);
  // a request whose intermediates exceed a worker's workspace can never run
  if (batch_workspace_bytes(s->workers[0].ctx, db, 1) > s->ws_bytes || r.footprint > s->slab.size)
    return XPG_ESIZE;
  std::lock_guard<std::mutex> lk(s->mu);
  uint64_t off = 0;
  // admission: the reply slot must fit the slots now free
  if (!s->slab.alloc(r.reply_words * 8, &off)) return XPG_EBUSY;
  r.reply = (uint64_t *)(s->slab.base + off);
  r.ticket = s->next_ticket++;
  r.submit_us = now_us();
  s->budget_used += r.footprint;
  s->slots[r.ticket] = {off, r.reply_words * 8};
  s->footprint[r.ticket] = r.footprint;
  s->pending.push_back(r);
  *ticket = r.ticket;
  s->cv.notify_one();
  return XPG_OK;
}

int xpg_server_poll(xpg_server *s, xpg_completion *out, uint32_t max, uint32_t *n_out, int64_t timeout_us) {
  if (!s || !n_out || (!out && max)) return XPG_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->done.empty() && timeout_us > 0)
    s->done_cv.wait_for(lk, std::chrono::microseconds(timeout_us), [&] { return !s->done.empty(); });
  const uint32_t n = (uint32_t)std::min<size_t>(max, s->done.size());
  for (uint32_t i = 0; i < n; i++) out[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + n);
  *n_out = n;
  return XPG_OK;
}

int xpg_server_release(xpg_server *s, uint64_t ticket) {
  if (!s) return XPG_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->slots.find(ticket);
  if (it == s->slots.end()) return XPG_EINVAL;
  // a ticket still pending or running keeps its slot: kernels may still write it
  if (!s->finished.count(ticket)) return XPG_EBUSY;
  s->finished.erase(ticket);
  s->slab.release(it->second.first, it->second.second);
  s->slots.erase(it);
  auto f = s->footprint.find(ticket);
  s->budget_used -= f->second;
  s->footprint.erase(f);
  return XPG_OK;
}

double xpg_server_now_us(xpg_server *) { return now_us(); }

int xpg_server_stats(xpg_server *s, uint64_t *batches, uint64_t *batched_requests, uint64_t *budget_used) {
  if (!s) return XPG_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  if (batches) *batches = s->batches;
  if (batched_requests) *batched_requests = s->batched_requests;
  if (budget_used) *budget_used = s->budget_used;
  return XPG_OK;
}

}  // extern "C"
