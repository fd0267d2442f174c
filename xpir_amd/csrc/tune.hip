// tune.hip -- the MAC autotuner and the per-device tuning table.
//
// Every MAC launch shape (one query's reply level, or a batch of queries on level 0) has
// several kernels that all give the same words: one chain per lane, register blocks over
// groups and queries, streaming block summaries, each
// with tile parameters.  Which is fastest depends on the shape (chain length, number of
// groups sharing a query word, polys, batch size).  xpg_autotune times every candidate on
// the database's real level shapes, checks that each candidate's words equal the untuned
// choice's, and records the fastest per shape in a table that every context on the
// device consults before falling back to the fixed rules.  No reference counterpart: the
// reference runs one CPU loop nest for every shape
// (pir/replyGenerator/PIRReplyGeneratorNFL_internal.cpp:511-595).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <vector>

#include "internal.hpp"

namespace {

struct TuneVal {
  LevelCfg cfg;
  float ms = 0.f, default_ms = 0.f;
};

std::mutex g_mu;
std::map<int, std::map<TuneKey, TuneVal>> g_table;  // device -> shape -> choice
std::atomic<uint64_t> g_gen{0};
std::atomic<uint64_t> g_size{0};

__global__ void words_differ_kernel(const uint64_t *a, const uint64_t *b, uint64_t n, uint64_t *flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    if (a[i] != b[i]) flag[0] = 1;
}

uint64_t prod_from(const xpg_db *db, uint32_t from) {
  uint64_t r = 1;
  for (uint32_t i = from; i < db->d; i++) r *= db->dims[i];
  return r;
}

// Device temporaries of one tuning run, freed on every exit path.
struct Temps {
  hipStream_t stream;
  std::vector<void *> ptrs;
  explicit Temps(hipStream_t s) : stream(s) {}
  ~Temps() {
    (void)hipStreamSynchronize(stream);  // nothing queued may still use them
    for (void *p : ptrs) (void)hipFree(p);
  }
  uint64_t *alloc(xpg_ctx *ctx, uint64_t words) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<uint64_t>(words, 1) * 8) != hipSuccess) {
      set_err(ctx, XPG_ENOMEM, "autotune temporaries");
      return nullptr;
    }
    ptrs.push_back(p);
    return (uint64_t *)p;
  }
};

struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
  Timer() {
    (void)hipEventCreateWithFlags(&a, hipEventDisableSystemFence);  // timing only
    (void)hipEventCreateWithFlags(&b, hipEventDisableSystemFence);
  }
  ~Timer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

// ms per launch of `run`: one timed launch, and when that took under 10 ms, enough
// back-to-back launches for ~20 ms (at most 10); negative on a launch error.  The caller
// runs the shape once untimed first (buffers sized, code loaded).
template <class F>
float time_launches(xpg_ctx *ctx, Timer &tm, F run) {
  if (hipEventRecord(tm.a, ctx->stream) != hipSuccess) return -1.f;
  if (run()) return -1.f;
  if (hipEventRecord(tm.b, ctx->stream) != hipSuccess || hipEventSynchronize(tm.b) != hipSuccess) return -1.f;
  float one = 0.f;
  if (hipEventElapsedTime(&one, tm.a, tm.b) != hipSuccess) return -1.f;
  if (one >= 10.f) return one;
  const int reps = std::max(1, std::min(10, (int)(20.f / std::max(one, 1e-3f))));
  (void)hipEventRecord(tm.a, ctx->stream);
  for (int i = 0; i < reps; i++)
    if (run()) return -1.f;
  (void)hipEventRecord(tm.b, ctx->stream);
  if (hipEventSynchronize(tm.b) != hipSuccess) return -1.f;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, tm.a, tm.b);
  return ms / reps;
}

// Runs `run` until ~30 ms of GPU time have passed (at least twice, at most 64 times), so
// that the first timed candidate does not pay for cold buffers or clocks.
template <class F>
int warm_up(xpg_ctx *ctx, Timer &tm, F run) {
  float spent = 0.f;
  for (int i = 0; i < 64 && (i < 2 || spent < 30.f); i++) {
    if (hipEventRecord(tm.a, ctx->stream) != hipSuccess) return XPG_EDEVICE;
    const int rc = run();
    if (rc) return rc;
    if (hipEventRecord(tm.b, ctx->stream) != hipSuccess || hipEventSynchronize(tm.b) != hipSuccess)
      return XPG_EDEVICE;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, tm.a, tm.b);
    spent += ms;
  }
  return XPG_OK;
}

// 1 if the two device word arrays differ, 0 if equal, negative on error.
int words_differ(xpg_ctx *ctx, const uint64_t *a, const uint64_t *b, uint64_t n, uint64_t *flag) {
  if (launch_fill_zero(ctx, flag, 1)) return -1;
  const uint64_t blocks = std::min<uint64_t>(4096, std::max<uint64_t>(1, (n + 255) / 256));
  hipLaunchKernelGGL(words_differ_kernel, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, a, b, n, flag);
  if (hipGetLastError() != hipSuccess) return -1;
  uint64_t h = 0;
  if (hipMemcpyAsync(&h, flag, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    return -1;
  return h ? 1 : 0;
}

bool same_cfg(const LevelCfg &x, const LevelCfg &y) {
  return x.kernel == y.kernel && x.a == y.a && x.b == y.b && x.c == y.c;
}

const uint64_t kScratchCap = 2ull << 30;  // candidates needing more scratch are skipped
const float kKeep = 0.97f;                // a candidate must beat the fixed rule by 3 %

std::vector<LevelCfg> level_candidates(const xpg_ctx *ctx, const xpg_params *prm, uint64_t n_l, uint64_t groups,
                                       uint32_t P, const LevelCfg &def) {
  std::vector<LevelCfg> v{def};
  auto add = [&](LevelCfg c) {
    for (const LevelCfg &x : v)
      if (same_cfg(x, c)) return;
    if (level_scratch_bytes(ctx, prm, n_l, groups, P, c) > kScratchCap) return;
    v.push_back(c);
  };
  add(LevelCfg{1, 0, 0, 0});
  if (groups >= 2) {
    for (int fam : {2, 3}) {
      add(LevelCfg{4, 4, 1, fam});
      add(LevelCfg{4, 8, 1, fam});
    }
    if (groups >= 16) add(LevelCfg{4, 16, 1, 2});
  }
  if (prm->n >= 64 && n_l >= 2) {
    for (int u : {4, 8, 16})
      for (int pf : {0, 1, 2}) {
        const uint64_t e0 = stream_auto_e(ctx, prm, n_l, groups, P, u);
        // one pass at four, two, one and eight waves per SIMD, and at three (the
        // register-heavier schedules fit three)
        for (uint64_t e : {e0, 2 * e0, 4 * e0, std::max<uint64_t>(u, e0 / 2), (e0 * 4 / 3 + u - 1) / u * u})
          if (e <= (1u << 20)) add(LevelCfg{3, u, pf, (int)e});
      }
  }
  return v;
}

void install(int device, const std::vector<xpg_tune_entry> &es) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto &t = g_table[device];
  for (const xpg_tune_entry &e : es) {
    const TuneKey k{e.kind, e.n, e.m, e.P, e.n_l, e.groups, e.qn};
    TuneVal v;
    v.cfg = LevelCfg{e.kernel, e.a, e.b, e.c};
    v.ms = e.ms;
    v.default_ms = e.default_ms;
    t[k] = v;
  }
  uint64_t s = 0;
  for (auto &kv : g_table) s += kv.second.size();
  g_size = s;
  g_gen++;
}

bool entry_valid(const xpg_tune_entry &e) {
  if (e.n < 64 || (e.n & (e.n - 1)) || e.m < 1 || e.m > XPG_MAX_M || e.P < 1 || e.n_l < 1 || e.groups < 1)
    return false;
  if (e.kind == 1)
    return e.qn >= 1 && e.qn <= 64 &&
           (e.kernel == 0 || (e.kernel == 4 && (e.c == 0 ? batch_shape_built(e.a, e.b)
                                                          : batch_family_built(e.c, e.a, e.b))));
  if (e.kind != 0 || e.qn != 0) return false;
  switch (e.kernel) {
    case 1: return true;
    case 3: return (e.a == 4 || e.a == 8 || e.a == 16) && e.b >= 0 && e.b <= 2 && e.c >= 0 && e.c <= (1 << 20);
    case 4: return (e.a == 0 || e.a == 1 || e.a == 4 || e.a == 8 || e.a == 16) && e.b == 1 &&
                   (e.c == 0 || e.c == 2 || e.c == 3) && (e.c != 3 || e.a == 4 || e.a == 8) && (e.a != 16 || e.c != 3);
    default: return false;
  }
}

xpg_tune_entry make_entry(uint32_t kind, const xpg_params *prm, uint32_t P, uint64_t n_l, uint64_t groups, uint32_t qn,
                          const LevelCfg &c, float ms, float def_ms) {
  xpg_tune_entry e{};
  e.kind = kind;
  e.n = prm->n;
  e.m = prm->m;
  e.P = P;
  e.n_l = n_l;
  e.groups = groups;
  e.qn = qn;
  e.kernel = c.kernel;
  e.a = c.a;
  e.b = c.b;
  e.c = c.c;
  e.ms = ms;
  e.default_ms = def_ms;
  return e;
}

// Restores the context's timing switch when the tuner returns.
struct TimingOff {
  xpg_ctx *ctx;
  bool was;
  explicit TimingOff(xpg_ctx *c) : ctx(c), was(c->timing) { ctx->timing = false; }
  ~TimingOff() { ctx->timing = was; }
};

}  // namespace

bool tune_lookup(int device, const TuneKey &k, LevelCfg *cfg) {
  if (g_size.load(std::memory_order_relaxed) == 0) return false;
  std::lock_guard<std::mutex> lk(g_mu);
  auto d = g_table.find(device);
  if (d == g_table.end()) return false;
  auto it = d->second.find(k);
  if (it == d->second.end()) return false;
  *cfg = it->second.cfg;
  return true;
}

uint64_t tune_generation() { return g_gen.load(); }

uint32_t tune_qn_bucket(uint32_t qn) {
  uint32_t b = 1;
  while (b < qn) b <<= 1;
  return b;
}

extern "C" {

int xpg_tune_set(xpg_ctx *ctx, const xpg_tune_entry *entries, uint32_t count) {
  if (!ctx || (!entries && count)) return XPG_EINVAL;
  std::vector<xpg_tune_entry> es(entries, entries + count);
  for (const xpg_tune_entry &e : es)
    if (!entry_valid(e)) return set_err(ctx, XPG_EINVAL, "tuning entry names a kernel or shape that is not built");
  install(ctx->device, es);
  return XPG_OK;
}

int xpg_tune_get(xpg_ctx *ctx, xpg_tune_entry *out, uint32_t cap, uint32_t *count) {
  if (!ctx || !count || (!out && cap)) return XPG_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  uint32_t i = 0;
  auto d = g_table.find(ctx->device);
  if (d != g_table.end()) {
    for (auto &kv : d->second) {
      if (i < cap) {
        const TuneKey &k = kv.first;
        xpg_tune_entry &e = out[i];
        e = xpg_tune_entry{};
        e.kind = k.kind;
        e.n = k.n;
        e.m = k.m;
        e.P = k.P;
        e.n_l = k.n_l;
        e.groups = k.groups;
        e.qn = k.qn;
        e.kernel = kv.second.cfg.kernel;
        e.a = kv.second.cfg.a;
        e.b = kv.second.cfg.b;
        e.c = kv.second.cfg.c;
        e.ms = kv.second.ms;
        e.default_ms = kv.second.default_ms;
      }
      i++;
    }
  }
  *count = i;
  return XPG_OK;
}

int xpg_tune_clear(xpg_ctx *ctx) {
  if (!ctx) return XPG_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  g_table.erase(ctx->device);
  uint64_t s = 0;
  for (auto &kv : g_table) s += kv.second.size();
  g_size = s;
  g_gen++;
  return XPG_OK;
}

int xpg_autotune(xpg_ctx *ctx, const xpg_db *db, uint32_t max_batch, xpg_tune_entry *out, uint32_t cap,
                 uint32_t *count) {
  if (!ctx || !db || (!out && cap)) return XPG_EINVAL;
  if (db->device != ctx->device) return set_err(ctx, XPG_EINVAL, "database lives on another GPU");
  if (max_batch > 64) return set_err(ctx, XPG_EINVAL, "batches hold at most 64 queries");
  if (db->mx) return set_err(ctx, XPG_EINVAL, "a database in the matrix layout has one level-0 kernel: nothing to tune");
  (void)hipSetDevice(ctx->device);
  TimingOff quiet(ctx);
  const xpg_params *prm = &db->prm;
  const uint64_t mn = (uint64_t)prm->m * prm->n;
  uint64_t qwords = 0;
  for (uint32_t l = 0; l < db->d; l++) qwords += db->dims[l];
  qwords *= 2 * mn;
  Temps tmp(ctx->stream);
  Timer tm;
  // a real query (index 0) under a fixed key: the chains see words of the right ranges
  std::vector<uint64_t> sk(mn);
  if (xpg_keygen(prm, 0x7475'6e65ull, sk.data()) != XPG_OK) return set_err(ctx, XPG_EINVAL, "keygen");
  uint64_t *skd = tmp.alloc(ctx, mn);
  uint64_t *flag = tmp.alloc(ctx, 1);
  uint64_t *qd = tmp.alloc(ctx, qwords);
  if (!skd || !flag || !qd) return XPG_ENOMEM;
  if (hipMemcpyAsync(skd, sk.data(), mn * 8, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
    return set_err(ctx, XPG_EDEVICE, "H2D copy");
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return set_err(ctx, XPG_EDEVICE, "stream sync");
  int rc = xpg_query_gen_dev(ctx, prm, skd, db->d, db->dims, 0, 1, qd);
  if (rc) return rc;
  std::vector<xpg_tune_entry> found;

  // ---- one query, every level ----
  float level0_ms = 0.f;
  const uint64_t *data = db->words;
  const uint64_t *q = qd;
  for (uint32_t l = 0; l < db->d; l++) {
    const uint64_t n_l = db->dims[l], groups = prod_from(db, l + 1);
    const uint32_t P = db->Pl[l];
    const uint64_t ow = groups * P * 2 * mn;
    uint64_t *ref = tmp.alloc(ctx, ow), *cand = tmp.alloc(ctx, ow);
    if (!ref || !cand) return XPG_ENOMEM;
    const LevelCfg def = default_level_cfg(ctx, prm, n_l, groups, P);
    const std::vector<LevelCfg> cands = level_candidates(ctx, prm, n_l, groups, P, def);
    float best_ms = -1.f, def_ms = -1.f;
    LevelCfg best = def;
    // untimed warm-up of the fixed rule's choice (buffers sized, clocks up), ~30 ms
    rc = warm_up(ctx, tm, [&] { return launch_mac_level_cfg(ctx, prm, data, q, n_l, groups, P, ref, def); });
    if (rc) return rc;
    for (size_t i = 0; i < cands.size(); i++) {
      uint64_t *dst = i == 0 ? ref : cand;
      const LevelCfg c = cands[i];
      const float ms = time_launches(ctx, tm, [&] {
        return launch_mac_level_cfg(ctx, prm, data, q, n_l, groups, P, dst, c);
      });
      if (ms < 0) {
        if (i == 0) return ctx->err.empty() ? set_err(ctx, XPG_EDEVICE, "autotune launch") : XPG_EDEVICE;
        (void)hipGetLastError();
        continue;  // a candidate this shape cannot take (grid or scratch limits)
      }
      if (i > 0) {
        const int d = words_differ(ctx, ref, cand, ow, flag);
        if (d < 0) return set_err(ctx, XPG_EDEVICE, "autotune compare");
        if (d) return set_err(ctx, XPG_EDEVICE, "autotune: a MAC kernel disagrees with the default choice");
      } else {
        def_ms = ms;
      }
      if (best_ms < 0 || ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
    if (!same_cfg(best, def) && best_ms <= kKeep * def_ms) {
      // confirm against noise: time the rule's choice and the winner again, back to back
      const float d2 = time_launches(ctx, tm, [&] {
        return launch_mac_level_cfg(ctx, prm, data, q, n_l, groups, P, cand, def);
      });
      const float b2 = time_launches(ctx, tm, [&] {
        return launch_mac_level_cfg(ctx, prm, data, q, n_l, groups, P, cand, best);
      });
      if (d2 > 0 && b2 > 0) {
        def_ms = std::min(def_ms, d2);
        best_ms = std::min(best_ms, b2);
      }
    }
    if (best_ms > kKeep * def_ms) {  // a tie within timing noise keeps the fixed rule
      best = def;
      best_ms = def_ms;
    }
    if (l == 0) level0_ms = best_ms;
    found.push_back(make_entry(0, prm, P, n_l, groups, 0, best, best_ms, def_ms));
    q += n_l * 2 * mn;
    if (l + 1 < db->d) {
      // fromResulttoInData on the level's reply: the next level's database
      const uint32_t P2 = db->Pl[l + 1];
      uint64_t *next = tmp.alloc(ctx, groups * P2 * mn);
      if (!next) return XPG_ENOMEM;
      const uint64_t unit = (uint64_t)P * 2 * mn * 8;
      ByteSrc src{(const uint8_t *)ref, unit, unit, groups * unit};
      rc = launch_import_bytes(ctx, prm, src, groups, groups, P2, next);
      if (rc) return rc;
      data = next;
    }
  }

  // ---- batches of queries on level 0 ----
  const uint64_t n0 = db->dims[0], groups0 = prod_from(db, 1);
  const uint32_t P0 = db->P0;
  const uint64_t q0w = n0 * 2 * mn, o0w = groups0 * P0 * 2 * mn;
  const uint64_t budget = 4ull << 30;  // device bytes for the batch's queries and outputs
  std::vector<uint64_t *> qs;          // distinct level-0 queries (level 0 part of full queries)
  for (uint32_t qn = 2; qn <= max_batch; qn *= 2) {
    if ((qn + 1) * o0w * 8 > budget) break;
    const uint64_t distinct = std::min<uint64_t>(qn, std::max<uint64_t>(1, (budget / 2) / (q0w * 8)));
    while (qs.size() < distinct) {
      uint64_t *b = tmp.alloc(ctx, qwords);
      if (!b) return XPG_ENOMEM;
      rc = xpg_query_gen_dev(ctx, prm, skd, db->d, db->dims, qs.size() % n0, 100 + qs.size(), b);
      if (rc) return rc;
      qs.push_back(b);
    }
    std::vector<const uint64_t *> qp(qn);
    for (uint32_t i = 0; i < qn; i++) qp[i] = qs[i % qs.size()];
    uint64_t *ref = tmp.alloc(ctx, qn * o0w), *cand = tmp.alloc(ctx, qn * o0w);
    if (!ref || !cand) return XPG_ENOMEM;
    std::vector<uint64_t *> rp(qn), cp(qn);
    for (uint32_t i = 0; i < qn; i++) {
      rp[i] = ref + i * o0w;
      cp[i] = cand + i * o0w;
    }
    // the shape picked by the fixed rules first (its family by shape: 0), then every other
    // built shape that fits, in both families
    int dg = 0, dq = 0;
    batch_default_shape(qn, groups0, &dg, &dq);
    struct Shape {
      int gr, qr, fam;
    };
    std::vector<Shape> shapes{{dg, dq, 0}};
    const int built[][2] = {{1, 1}, {1, 2}, {1, 4}, {4, 1}, {4, 2}, {4, 4}, {8, 1}, {8, 2}, {16, 1}};
    const int def_fam = batch3_shape_built(dg, dq) ? 3 : 2;
    for (const auto &s : built)
      for (int fam : {2, 3}) {
        const bool ok = batch_family_built(fam, s[0], s[1]);
        if (ok && (s[0] != dg || s[1] != dq || fam != def_fam) && s[1] <= (int)qn &&
            (s[0] == 1 || groups0 * P0 >= (uint64_t)s[0] / 2))
          shapes.push_back(Shape{s[0], s[1], fam});
      }
    rc = warm_up(ctx, tm, [&] {
      return launch_mac_batch_cfg(ctx, prm, db->words, qp.data(), rp.data(), qn, n0, groups0, P0, dg, dq, 0);
    });
    if (rc) return rc;
    float best_ms = -1.f, def_ms = -1.f;
    int bg = 0, bq = 0, bf = 0;
    for (size_t i = 0; i < shapes.size(); i++) {
      const int gr = shapes[i].gr, qr = shapes[i].qr, fam = shapes[i].fam;
      std::vector<uint64_t *> &dst = i == 0 ? rp : cp;
      const float ms = time_launches(ctx, tm, [&] {
        return launch_mac_batch_cfg(ctx, prm, db->words, qp.data(), dst.data(), qn, n0, groups0, P0, gr, qr, fam);
      });
      if (ms < 0) {
        if (i == 0) return ctx->err.empty() ? set_err(ctx, XPG_EDEVICE, "autotune launch") : XPG_EDEVICE;
        (void)hipGetLastError();
        continue;
      }
      if (i > 0) {
        const int d = words_differ(ctx, ref, cand, qn * o0w, flag);
        if (d < 0) return set_err(ctx, XPG_EDEVICE, "autotune compare");
        if (d) return set_err(ctx, XPG_EDEVICE, "autotune: a batched MAC shape disagrees with the default");
      } else {
        def_ms = ms;
      }
      if (best_ms < 0 || ms < best_ms) {
        best_ms = ms;
        bg = gr;
        bq = qr;
        bf = fam;
      }
    }
    if ((bg != dg || bq != dq || bf != 0) && best_ms <= kKeep * def_ms) {
      const float d2 = time_launches(ctx, tm, [&] {
        return launch_mac_batch_cfg(ctx, prm, db->words, qp.data(), cp.data(), qn, n0, groups0, P0, dg, dq, 0);
      });
      const float b2 = time_launches(ctx, tm, [&] {
        return launch_mac_batch_cfg(ctx, prm, db->words, qp.data(), cp.data(), qn, n0, groups0, P0, bg, bq, bf);
      });
      if (d2 > 0 && b2 > 0) {
        def_ms = std::min(def_ms, d2);
        best_ms = std::min(best_ms, b2);
      }
    }
    if (best_ms > kKeep * def_ms) {
      bg = dg;
      bq = dq;
      bf = 0;
      best_ms = def_ms;
    }
    // kernel 0: the batch's level 0 runs faster as qn single-query launches
    const LevelCfg choice = best_ms > (float)qn * level0_ms ? LevelCfg{0, 0, 0, 0} : LevelCfg{4, bg, bq, bf};
    found.push_back(make_entry(1, prm, P0, n0, groups0, qn, choice, std::min(best_ms, (float)qn * level0_ms),
                               def_ms));
  }
  install(ctx->device, found);
  if (out)
    for (uint32_t i = 0; i < std::min<uint32_t>(cap, (uint32_t)found.size()); i++) out[i] = found[i];
  if (count) *count = (uint32_t)found.size();
  return XPG_OK;
}

}  // extern "C"
