// capi.hip -- host side of libxpir_amd.so: parameters, NTT tables, contexts, resident
// databases and the d-level reply orchestration behind the C ABI (include/xpir_amd.h).
#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>
#include <vector>

#include "internal.hpp"

static std::atomic<uint64_t> g_db_serial{0};

// ---- scratch of the no-allocation reply (xpg_reply_dev_noalloc) ----
// That entry point takes its level buffers and the streaming MAC's block summaries from
// this array instead of hipMalloc.  Every GPU the library runs on holds its own copy
// until the process ends.  In host code the name `scratch` is only a stand-in: a
// device's copy is reached through the address hipGetSymbolAddress returns with that
// device current, looked up once per device when the first context on it is created.
static __device__ __align__(256) uint8_t scratch[XPG_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// With `device` current: its copy's address, once.  A failed lookup is kept and reported
// by xpg_reply_dev_noalloc; nothing else depends on the array.
void map_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess && p) {
    g_scratch_ptr[device] = static_cast<uint8_t *>(p);
    g_scratch_err[device] = hipSuccess;
  } else {
    g_scratch_err[device] = e == hipSuccess ? hipErrorInvalidSymbol : e;
  }
}

uint8_t *scratch_of(int device, hipError_t *err) {
  if (device < 0 || device >= kMaxDevices) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  *err = g_scratch_err[device];
  return g_scratch_ptr[device];
}
}  // namespace

namespace {

using u128 = unsigned __int128;

// Moduli and primitive 2^14-th roots of unity fixed by the reference
// (crypto/NFLParams.cpp:34,41).  The roots decide every NTT-domain word.
const uint64_t kModuli[4] = {2305843009213317121ULL, 2305843009213120513ULL, 2305843009212694529ULL,
                             2305843009212399617ULL};
const uint64_t kRoots[4] = {1187132827279672845ULL, 753478288701480417ULL, 717492273700781398ULL,
                            1965831349000139179ULL};

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)(((u128)a * b) % p); }
uint64_t quot(uint64_t y, uint64_t p) { return (uint64_t)(((u128)y << 64) / p); }

#define HIP_OK(expr, msg)                                                      \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return set_err(ctx, XPG_EDEVICE, std::string(msg) + ": " + hipGetErrorString(e_)); \
  } while (0)

uint64_t prod_dims(uint32_t d, const uint32_t *dims, uint32_t from) {
  uint64_t r = 1;
  for (uint32_t i = from; i < d; i++) r *= dims[i];
  return r;
}

}  // namespace

int set_err(xpg_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

ModSet modset_of(const xpg_params *prm) {
  ModSet ms{};
  for (uint32_t c = 0; c < prm->m && c < XPG_MAX_M; c++) {
    ms.p[c] = prm->moduli[c];
    ms.delta[c] = (uint32_t)((1ULL << 61) - prm->moduli[c]);
  }
  return ms;
}

// configureNTT + prep_wtab restated (crypto/NFLlib.cpp:44-194, NFLlib.hpp:571-586)
const DevTables *get_tables(xpg_ctx *ctx, uint32_t n, uint32_t m) {
  const uint64_t key = ((uint64_t)n << 8) | m;
  auto it = ctx->tables.find(key);
  if (it != ctx->tables.end()) return &it->second;
  if (ctx->noalloc) {
    set_err(ctx, XPG_EINVAL, "no NTT tables for these parameters on this context (the no-allocation reply builds none)");
    return nullptr;
  }
  uint32_t logn = 0;
  while ((1u << logn) < n) logn++;
  std::vector<uint64_t> phi(n * m), phiq(n * m), w(n * m), wq(n * m), iw(n * m), iwq(n * m), ipt(n * m),
      iptq(n * m);
  // twiddles layer after layer: sizes n/2, n/4, ..., 1 holding om^(i * 2^L)
  auto layers = [&](std::vector<uint64_t> &tab, std::vector<uint64_t> &tabq, size_t o, uint64_t om, uint64_t p) {
    for (uint32_t size = n; size >= 2; size >>= 1) {
      uint64_t wi = 1;
      for (uint32_t i = 0; i < size / 2; i++) {
        tab[o] = wi;
        tabq[o] = quot(wi, p);
        o++;
        wi = mulmod(wi, om, p);
      }
      om = mulmod(om, om, p);
    }
  };
  for (uint32_t c = 0; c < m; c++) {
    const uint64_t p = kModuli[c];
    uint64_t f = kRoots[c];
    for (uint32_t i = 0; i < 13 - logn; i++) f = mulmod(f, f, p);  // primitive 2n-th root
    uint64_t v = 1;
    for (uint32_t j = 0; j < n; j++) {
      phi[c * n + j] = v;
      phiq[c * n + j] = quot(v, p);
      v = mulmod(v, f, p);
    }
    // v = phi^n = -1, so phi^-1 = -phi^(n-1); n^-1 = p - (p-1)/n since 2n divides p - 1
    const uint64_t fi = mulmod(v, phi[c * n + n - 1], p);
    uint64_t u = p - (p - 1) / n;
    for (uint32_t j = 0; j < n; j++) {
      ipt[c * n + j] = u;
      iptq[c * n + j] = quot(u, p);
      u = mulmod(u, fi, p);
    }
    layers(w, wq, c * n, mulmod(f, f, p), p);
    layers(iw, iwq, c * n, mulmod(fi, fi, p), p);
  }
  DevTables t;
  t.n = n;
  t.m = m;
  t.logn = logn;
  const size_t bytes = sizeof(uint64_t) * n * m;
  uint64_t **dst[8] = {&t.phi, &t.phiq, &t.w, &t.wq, &t.iw, &t.iwq, &t.ipt, &t.iptq};
  const std::vector<uint64_t> *src[8] = {&phi, &phiq, &w, &wq, &iw, &iwq, &ipt, &iptq};
  for (int i = 0; i < 8; i++) {
    if (hipMalloc((void **)dst[i], bytes) != hipSuccess ||
        hipMemcpy(*dst[i], src[i]->data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
      for (int j = 0; j <= i; j++)
        if (*dst[j]) (void)hipFree(*dst[j]);
      set_err(ctx, XPG_ENOMEM, "NTT table allocation");
      return nullptr;
    }
  }
  return &(ctx->tables[key] = t);
}

int ensure_level_buf(xpg_ctx *ctx, int which, size_t bytes) {
  if (ctx->level_bytes[which] >= bytes) return XPG_OK;
  if (ctx->level_buf[which]) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->level_buf[which]);
  }
  ctx->level_buf[which] = nullptr;
  ctx->level_bytes[which] = 0;
  ctx->epoch++;
  HIP_OK(hipMalloc((void **)&ctx->level_buf[which], bytes), "level buffer");
  ctx->level_bytes[which] = bytes;
  return XPG_OK;
}

// Host memcpy split over threads (a pageable source has to pass through the CPU; one
// thread copies at a fraction of the host's memory bandwidth).
static void copy_host_threads(void *dst, const void *src, size_t bytes) {
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t T = std::min<size_t>(hw ? std::min(hw, 8u) : 4, std::max<size_t>(1, bytes >> 22));  // >= 4 MiB each
  if (T <= 1) {
    memcpy(dst, src, bytes);
    return;
  }
  const size_t per = (bytes + T - 1) / T;
  std::vector<std::thread> th;
  for (size_t i = 0; i < T && i * per < bytes; i++) {
    const size_t o = i * per, len = std::min(per, bytes - o);
    th.emplace_back([=] { memcpy((char *)dst + o, (const char *)src + o, len); });
  }
  for (auto &t : th) t.join();
}

// Host -> device copy of a pageable buffer: up to 64 MiB as one hipMemcpyAsync, larger ones
// in 64 MiB chunks through two pinned host chunks, host threads filling one while the DMA
// engine drains the other.  The source may be reused on return (every byte has been read
// into a pinned chunk or by the HIP runtime's own staging).
static constexpr size_t kPinChunk = 64ull << 20;

int copy_h2d_staged(xpg_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (bytes <= kPinChunk) {
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "H2D copy");
    return XPG_OK;
  }
  for (int b = 0; b < 2; b++) {
    if (!ctx->pin[b]) HIP_OK(hipHostMalloc(&ctx->pin[b], kPinChunk, hipHostMallocDefault), "pinned staging chunk");
    if (!ctx->pin_ev[b]) HIP_OK(hipEventCreateWithFlags(&ctx->pin_ev[b], hipEventDisableTiming), "staging event");
  }
  size_t i = 0;
  for (size_t off = 0; off < bytes; off += kPinChunk, i++) {
    const int b = (int)(i & 1);
    const size_t len = std::min(kPinChunk, bytes - off);
    HIP_OK(hipEventSynchronize(ctx->pin_ev[b]), "staging wait");  // its last copy has read it
    copy_host_threads(ctx->pin[b], (const char *)src + off, len);
    HIP_OK(hipMemcpyAsync((char *)dst + off, ctx->pin[b], len, hipMemcpyHostToDevice, ctx->stream), "H2D copy");
    HIP_OK(hipEventRecord(ctx->pin_ev[b], ctx->stream), "staging event");
  }
  return XPG_OK;
}

int copy_h2d_table(xpg_ctx *ctx, void *dst, const void *src, size_t bytes) {
  const int b = ctx->tab_next;
  ctx->tab_next ^= 1;
  if (!ctx->tab_ev[b]) HIP_OK(hipEventCreateWithFlags(&ctx->tab_ev[b], hipEventDisableTiming), "table event");
  HIP_OK(hipEventSynchronize(ctx->tab_ev[b]), "table wait");  // its last copy has read it
  if (ctx->tab_bytes[b] < bytes) {
    if (ctx->tab_pin[b]) (void)hipHostFree(ctx->tab_pin[b]);
    ctx->tab_pin[b] = nullptr;
    ctx->tab_bytes[b] = 0;
    const size_t want = std::max<size_t>(bytes, 64 << 10);
    HIP_OK(hipHostMalloc(&ctx->tab_pin[b], want, hipHostMallocDefault), "pinned table buffer");
    ctx->tab_bytes[b] = want;
  }
  memcpy(ctx->tab_pin[b], src, bytes);
  HIP_OK(hipMemcpyAsync(dst, ctx->tab_pin[b], bytes, hipMemcpyHostToDevice, ctx->stream), "table copy");
  HIP_OK(hipEventRecord(ctx->tab_ev[b], ctx->stream), "table event");
  return XPG_OK;
}

int ensure_io_buf(xpg_ctx *ctx, size_t bytes) {
  if (ctx->io_bytes >= bytes) return XPG_OK;
  if (ctx->io_buf) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->io_buf);
  }
  ctx->io_buf = nullptr;
  ctx->io_bytes = 0;
  ctx->epoch++;
  HIP_OK(hipMalloc((void **)&ctx->io_buf, bytes), "query staging buffer");
  ctx->io_bytes = bytes;
  return XPG_OK;
}

int ensure_dec_buf(xpg_ctx *ctx, size_t bytes) {
  if (ctx->dec_bytes >= bytes) return XPG_OK;
  if (ctx->dec_buf) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->dec_buf);
  }
  ctx->dec_buf = nullptr;
  ctx->dec_bytes = 0;
  ctx->epoch++;
  HIP_OK(hipMalloc((void **)&ctx->dec_buf, bytes), "decryption buffer");
  ctx->dec_bytes = bytes;
  return XPG_OK;
}

int ensure_batch_buf(xpg_ctx *ctx, size_t bytes) {
  if (ctx->batch_bytes >= bytes) return XPG_OK;
  if (ctx->batch_buf) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->batch_buf);
  }
  ctx->batch_buf = nullptr;
  ctx->batch_bytes = 0;
  ctx->epoch++;
  HIP_OK(hipMalloc((void **)&ctx->batch_buf, bytes), "batch buffer");
  ctx->batch_bytes = bytes;
  return XPG_OK;
}

int ensure_scratch(xpg_ctx *ctx, size_t bytes) {
  if (ctx->scratch_bytes >= bytes) return XPG_OK;
  if (ctx->scratch) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->scratch);
  }
  ctx->scratch = nullptr;
  ctx->scratch_bytes = 0;
  ctx->epoch++;
  HIP_OK(hipMalloc(&ctx->scratch, bytes), "scratch");
  ctx->scratch_bytes = bytes;
  return XPG_OK;
}

int get_scratch(xpg_ctx *ctx, size_t bytes, void **p) {
  if (ctx->ws) {
    if (bytes > ctx->ws_cap - ctx->ws_used)
      return set_err(ctx, XPG_ESIZE, "the workspace is too small for this launch's scratch");
    *p = ctx->ws + ctx->ws_used;  // per launch: later launches on the stream reuse it
    return XPG_OK;
  }
  const int rc = ensure_scratch(ctx, bytes);
  if (rc) return rc;
  *p = ctx->scratch;
  return XPG_OK;
}

void graph_cache_clear(xpg_ctx *ctx, uint64_t db_serial) {
  auto &c = ctx->graph_cache;
  for (size_t i = 0; i < c.size();) {
    if (db_serial == 0 || c[i].db_serial == db_serial) {
      if (c[i].exec) (void)hipGraphExecDestroy(c[i].exec);
      c[i] = c.back();
      c.pop_back();
    } else {
      i++;
    }
  }
}

void timing_begin(xpg_ctx *ctx, hipEvent_t *a, hipEvent_t *b, int tag) {
  // mode 2 times the main kernels alone: no events around whole levels, whose records
  // would put gaps between a level's launches; the no-allocation reply creates no events
  if (!ctx->timing || ctx->noalloc || (ctx->timing == 2 && tag == 0)) return;
  for (hipEvent_t *e : {a, b}) {
    if (!ctx->ev_pool.empty()) {
      *e = ctx->ev_pool.back();
      ctx->ev_pool.pop_back();
    } else {
      // timing only: no system-scope fence at the record (no cache writeback between the
      // launches it separates)
      (void)hipEventCreateWithFlags(e, hipEventDisableSystemFence);
    }
  }
  (void)hipEventRecord(*a, ctx->stream);
}

void timing_end(xpg_ctx *ctx, hipEvent_t a, hipEvent_t b, int tag) {
  if (!ctx->timing || !a) return;
  (void)hipEventRecord(b, ctx->stream);
  ctx->pending.emplace_back(a, b, tag);
}

static void timing_collect(xpg_ctx *ctx) {
  for (auto &pr : ctx->pending) {
    float ms = 0.f;
    hipEvent_t a = std::get<0>(pr), b = std::get<1>(pr);
    if (hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess) {
      if (std::get<2>(pr) != 0) {
        ctx->main_ms += ms;
        ctx->main_launches++;
        if (std::get<2>(pr) == 2) {
          ctx->batch_ms += ms;
          ctx->batch_launches++;
        }
      } else {
        ctx->mac_ms += ms;
        ctx->mac_launches++;
      }
    }
    ctx->ev_pool.push_back(a);
    ctx->ev_pool.push_back(b);
  }
  ctx->pending.clear();
}

// ------------------------------------------------------------------------------------
extern "C" {

int xpg_params_from_string(const char *desc, uint64_t n0, xpg_params *out) {
  if (!desc || !out) return XPG_EINVAL;
  unsigned k = 0, n = 0, bits = 0;
  char name[8] = {0};
  if (sscanf(desc, "%3[A-Z]:%u:%u:%u", name, &k, &n, &bits) != 4) return XPG_EINVAL;
  if (strcmp(name, "LWE") != 0) return XPG_EINVAL;
  if (bits == 0 || bits % 60 != 0 || bits / 60 > 4) return XPG_EINVAL;  // NFLlib.cpp:219-224
  if (n < 512 || n > 8192 || (n & (n - 1))) return XPG_EINVAL;
  memset(out, 0, sizeof(*out));
  out->k = k;
  out->n = n;
  while ((1u << out->logn) < n) out->logn++;
  out->m = bits / 60;
  out->bits = bits;
  out->Berr = 5 * k / 2;  // NFLLWE.cpp:84
  out->n0 = n0 ? n0 : 1;
  for (uint32_t c = 0; c < out->m; c++) out->moduli[c] = kModuli[c];
  // NFLLWE.cpp:656-668, in double precision as the reference evaluates it
  double a = floor(((double)(out->m * 60) - 1.0 - log2((double)out->n0) - log2((double)out->Berr) -
                    log2((double)n)) / 2.0);
  if (!(a >= 1.0) || a > 224.0) return XPG_EINVAL;
  out->A = (int32_t)a;
  out->abs_bytes = (uint64_t)out->A * n / 8;
  out->ct_bytes = 16ULL * n * out->m;
  return XPG_OK;
}

uint32_t xpg_polys_for_bytes(const xpg_params *prm, uint64_t unit_bytes) {
  const uint64_t per = (uint64_t)prm->A * prm->n;
  return (uint32_t)((unit_bytes * 8 + per - 1) / per);
}

uint32_t xpg_reply_polys(const xpg_params *prm, uint32_t P0, uint32_t d) {
  uint64_t P = P0;
  for (uint32_t l = 1; l < d; l++) P = (P * prm->ct_bytes + prm->abs_bytes - 1) / prm->abs_bytes;
  return (uint32_t)P;
}

int xpg_ctx_create(int device, xpg_ctx **out) {
  if (!out) return XPG_EINVAL;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return XPG_EDEVICE;
  if (hipSetDevice(device) != hipSuccess) return XPG_EDEVICE;
  xpg_ctx *ctx = new xpg_ctx();
  ctx->device = device;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    ctx->cu_count = cus;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return XPG_EDEVICE;
  }
  map_scratch(device);
  *out = ctx;
  return XPG_OK;
}

void xpg_ctx_destroy(xpg_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  timing_collect(ctx);
  for (xpg_db *db : ctx->dbs) db->ctx = nullptr;  // they stay resident until xpg_db_free
  for (auto &kv : ctx->tables) {
    for (uint64_t *b : {kv.second.phi, kv.second.phiq, kv.second.w, kv.second.wq, kv.second.iw, kv.second.iwq,
                        kv.second.ipt, kv.second.iptq})
      (void)hipFree(b);
  }
  for (int i = 0; i < 2; i++)
    if (ctx->level_buf[i]) (void)hipFree(ctx->level_buf[i]);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->batch_buf) (void)hipFree(ctx->batch_buf);
  if (ctx->fused_buf) (void)hipFree(ctx->fused_buf);
  if (ctx->io_buf) (void)hipFree(ctx->io_buf);
  for (int b = 0; b < 2; b++) {
    if (ctx->pin[b]) (void)hipHostFree(ctx->pin[b]);
    if (ctx->pin_ev[b]) (void)hipEventDestroy(ctx->pin_ev[b]);
    if (ctx->tab_pin[b]) (void)hipHostFree(ctx->tab_pin[b]);
    if (ctx->tab_ev[b]) (void)hipEventDestroy(ctx->tab_ev[b]);
  }
  if (ctx->dec_buf) (void)hipFree(ctx->dec_buf);
  for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
  graph_cache_clear(ctx, 0);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *xpg_last_error(const xpg_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }
void *xpg_ctx_stream(xpg_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

static int set_option(xpg_ctx *ctx, const char *key, int64_t value) {
  if (strcmp(key, "group_budget") == 0 && value >= 0) {
    ctx->group_budget = (uint64_t)value;
    return XPG_OK;
  }
  if (strcmp(key, "batch_fuse") == 0 && (value == 0 || value == 1)) {
    ctx->batch_fuse = (int)value;
    return XPG_OK;
  }
  if (strcmp(key, "alpha_zero_pad") == 0 && (value == 0 || value == 1)) {
    ctx->alpha_zero_pad = (int)value;
    return XPG_OK;
  }
  if (strcmp(key, "db_layout") == 0 && (value == 0 || value == 1)) {
    ctx->db_layout = (int)value;
    return XPG_OK;
  }
  if (strcmp(key, "graphs") == 0 && (value == 0 || value == 1)) {
    ctx->graphs = value != 0;
    if (!ctx->graphs) graph_cache_clear(ctx, 0);
    return XPG_OK;
  }
  if (strcmp(key, "mac_kernel") == 0 && (value == 0 || value == 1 || value == 3)) {
    ctx->mac_mode = (int)value;
    return XPG_OK;
  }
  if (strcmp(key, "stream_pf") == 0 && value >= 0 && value <= 3) {
    ctx->stream_pf = (int)value;
    return XPG_OK;
  }
  if (strcmp(key, "stream_u") == 0 && (value == 0 || value == 4 || value == 8 || value == 16)) {
    ctx->stream_u = (int)value;
    return XPG_OK;
  }
  if (strcmp(key, "stream_e") == 0 && value >= 0 && value <= (1 << 20)) {
    ctx->stream_e = (int)value;
    return XPG_OK;
  }
  return set_err(ctx, XPG_EINVAL, std::string("unknown option or value: ") + key);
}

int xpg_ctx_set_option(xpg_ctx *ctx, const char *key, int64_t value) {
  if (!ctx || !key) return XPG_EINVAL;
  const int rc = set_option(ctx, key, value);
  if (!rc) ctx->epoch++;  // captured replies baked the old settings in
  return rc;
}

int xpg_sync(xpg_ctx *ctx) {
  if (!ctx) return XPG_EINVAL;
  HIP_OK(hipStreamSynchronize(ctx->stream), "stream sync");
  return XPG_OK;
}

int xpg_dev_alloc(xpg_ctx *ctx, uint64_t bytes, void **ptr) {
  if (!ctx || !ptr) return XPG_EINVAL;
  (void)hipSetDevice(ctx->device);
  if (hipMalloc(ptr, bytes ? bytes : 8) != hipSuccess) return set_err(ctx, XPG_ENOMEM, "device allocation");
  return XPG_OK;
}

int xpg_dev_free(xpg_ctx *ctx, void *ptr) {
  if (!ctx) return XPG_EINVAL;
  HIP_OK(hipStreamSynchronize(ctx->stream), "stream sync");
  HIP_OK(hipFree(ptr), "device free");
  return XPG_OK;
}

int xpg_h2d(xpg_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx) return XPG_EINVAL;
  const int rc = copy_h2d_staged(ctx, dst, src, bytes);
  if (rc) return rc;
  HIP_OK(hipStreamSynchronize(ctx->stream), "stream sync");
  return XPG_OK;
}

int xpg_host_alloc(xpg_ctx *ctx, uint64_t bytes, void **ptr) {
  if (!ctx || !ptr) return XPG_EINVAL;
  (void)hipSetDevice(ctx->device);
  if (hipHostMalloc(ptr, bytes ? bytes : 8, hipHostMallocDefault) != hipSuccess) {
    *ptr = nullptr;
    return set_err(ctx, XPG_ENOMEM, "pinned host allocation");
  }
  return XPG_OK;
}

int xpg_host_free(xpg_ctx *ctx, void *ptr) {
  if (!ctx) return XPG_EINVAL;
  HIP_OK(hipStreamSynchronize(ctx->stream), "stream sync");  // no copy from it in flight
  HIP_OK(hipHostFree(ptr), "pinned host free");
  return XPG_OK;
}

int xpg_h2d_async(xpg_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx || !dst || !src) return XPG_EINVAL;
  HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "H2D copy");
  return XPG_OK;
}

int xpg_d2h(xpg_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx) return XPG_EINVAL;
  HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "D2H copy");
  HIP_OK(hipStreamSynchronize(ctx->stream), "stream sync");
  return XPG_OK;
}

// ---- databases ----
static int db_setup(xpg_ctx *ctx, const xpg_params *prm, uint64_t n_elems, uint64_t elem_bytes,
                    uint32_t alpha, uint32_t d, const uint32_t *dims, uint32_t P0, uint64_t real_units,
                    xpg_db **out) {
  if (!ctx || !prm || !out || !dims || d == 0 || d > XPG_MAX_LEVELS || alpha == 0)
    return set_err(ctx, XPG_EINVAL, "bad database arguments");
  for (uint32_t l = 0; l < d; l++)
    if (dims[l] == 0) return set_err(ctx, XPG_EINVAL, "zero dimension");
  if (dims[0] != prm->n0) return set_err(ctx, XPG_EINVAL, "params.n0 must equal dims[0] (A depends on it)");
  const uint64_t units = prod_dims(d, dims, 0);
  if (units < real_units) return set_err(ctx, XPG_EINVAL, "grid smaller than the database");
  xpg_db *db = new xpg_db();
  db->ctx = ctx;
  db->device = ctx->device;
  db->serial = ++g_db_serial;
  db->prm = *prm;
  db->d = d;
  for (uint32_t l = 0; l < d; l++) db->dims[l] = dims[l];
  db->units = units;
  db->real_units = real_units;
  db->P0 = P0;
  db->Pl[0] = P0;
  for (uint32_t l = 1; l <= d; l++) {
    db->Pl[l] = (l < d) ? (uint32_t)((db->Pl[l - 1] * prm->ct_bytes + prm->abs_bytes - 1) / prm->abs_bytes)
                        : db->Pl[l - 1];
  }
  db->n_elems = n_elems;
  db->elem_bytes = elem_bytes;
  db->alpha = alpha;
  db->alpha_zero_pad = ctx->alpha_zero_pad;
  const uint64_t words = units * P0 * (uint64_t)prm->m * prm->n;
  (void)hipSetDevice(ctx->device);
  if (hipMalloc((void **)&db->words, words * 8) != hipSuccess) {
    delete db;
    return set_err(ctx, XPG_ENOMEM, "NTT database allocation");
  }
  ctx->dbs.insert(db);
  *out = db;
  return XPG_OK;
}

// The last import unit when n_elems is not a multiple of alpha, with the bytes the
// reference's importDataNFL gives it (PIRReplyGeneratorNFL_internal.cpp:86-105 over
// DBHandler::readAggregatedStream, DBHandler.cpp:3-26): the buffer is shared by all
// units; the last unit's r real elements go to slots 0..r-1, then alpha - r slots from
// slot r-1 are zeroed, so the last real element is erased and slot alpha-1 keeps the
// previous unit's last element (zeros when there is no previous unit).  The unit is
// staged in a small device buffer and imported with the grid's padding units after it.
static int import_partial_aggregate(xpg_ctx *ctx, const xpg_params *prm, const uint8_t *elems_dev, uint64_t n_elems,
                                    uint64_t elem_bytes, uint32_t alpha, uint32_t P0, xpg_db *db) {
  const uint64_t r = n_elems % alpha, k = n_elems / alpha;  // k full units before the last
  const uint64_t unit_bytes = elem_bytes * alpha;
  uint8_t *tmp = nullptr;
  HIP_OK(hipMalloc((void **)&tmp, unit_bytes + 16), "aggregate staging");
  int rc = XPG_OK;
  const uint8_t *first = elems_dev + k * unit_bytes;
  if (r > 1 && hipMemcpyAsync(tmp, first, (r - 1) * elem_bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
    rc = set_err(ctx, XPG_EDEVICE, "aggregate staging copy");
  if (!rc && hipMemsetAsync(tmp + (r - 1) * elem_bytes, 0, (alpha - r) * elem_bytes, ctx->stream) != hipSuccess)
    rc = set_err(ctx, XPG_EDEVICE, "aggregate staging clear");
  uint8_t *last_slot = tmp + (alpha - 1) * elem_bytes;
  if (!rc && k > 0 &&
      hipMemcpyAsync(last_slot, elems_dev + (k * alpha - 1) * elem_bytes, elem_bytes, hipMemcpyDeviceToDevice,
                     ctx->stream) != hipSuccess)
    rc = set_err(ctx, XPG_EDEVICE, "aggregate staging copy");
  if (!rc && k == 0 && hipMemsetAsync(last_slot, 0, elem_bytes, ctx->stream) != hipSuccess)
    rc = set_err(ctx, XPG_EDEVICE, "aggregate staging clear");
  if (!rc) {
    ByteSrc src{tmp, unit_bytes, unit_bytes, unit_bytes};
    const uint64_t per_unit = (uint64_t)P0 * prm->m * prm->n;
    rc = launch_import_bytes(ctx, prm, src, 1, db->units - k, P0, db->words + k * per_unit);
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(tmp);
  return rc;
}

// The database in the matrix layout when the context asks for it (option db_layout) and
// its level 0 has one (mx_geometry): level 0's digit fragments and tail rows are made from
// the row-major words, which are then freed.  A database whose level 0 is too short keeps
// its row-major words.
static int db_to_matrix(xpg_ctx *ctx, xpg_db *db) {
  if (ctx->db_layout != 1) return XPG_OK;
  MxGeom g;
  if (!mx_geometry(&db->prm, db->dims[0], prod_dims(db->d, db->dims, 1), db->P0, &g)) return XPG_OK;
  uint32_t *bad = nullptr;
  HIP_OK(hipMalloc((void **)&db->mx, mx_frag_bytes(g)), "matrix layout allocation");
  HIP_OK(hipMalloc((void **)&db->mx_tail, mx_tail_words(g) * 8), "tail rows allocation");
  HIP_OK(hipMalloc((void **)&bad, 256), "flag allocation");
  int rc = XPG_OK;
  if (hipMemsetAsync(bad, 0, 4, ctx->stream) != hipSuccess) rc = set_err(ctx, XPG_EDEVICE, "flag clear");
  if (!rc) rc = launch_mx_pack_db(ctx, &db->prm, g, db->words, db->mx, db->mx_tail, bad);
  uint32_t hbad = 0;
  if (!rc && hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
    rc = set_err(ctx, XPG_EDEVICE, "flag copy");
  if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc) rc = set_err(ctx, XPG_EDEVICE, "stream sync");
  (void)hipFree(bad);
  if (!rc && hbad) rc = set_err(ctx, XPG_EINVAL, "a database word is too large for the matrix layout (>= 2^63 - 2^56)");
  if (rc) {
    (void)hipFree(db->mx);
    (void)hipFree(db->mx_tail);
    db->mx = nullptr;
    db->mx_tail = nullptr;
    return rc;
  }
  (void)hipFree(db->words);
  db->words = nullptr;
  db->mxg = g;
  return XPG_OK;
}

int xpg_db_import_dev(xpg_ctx *ctx, const xpg_params *prm, const uint8_t *elems_dev, uint64_t n_elems,
                      uint64_t elem_bytes, uint32_t alpha, uint32_t d, const uint32_t *dims, xpg_db **out) {
  if (!ctx || !prm || !out) return XPG_EINVAL;
  if (((uintptr_t)elems_dev & 7) != 0) return set_err(ctx, XPG_EINVAL, "element bytes must be 8-byte aligned");
  if (elem_bytes == 0 || n_elems == 0) return set_err(ctx, XPG_EINVAL, "empty database");
  if (alpha == 0) return set_err(ctx, XPG_EINVAL, "alpha must be at least 1");
  const uint64_t unit_bytes = elem_bytes * alpha;
  const uint32_t P0 = xpg_polys_for_bytes(prm, unit_bytes);
  const uint64_t real_units = (n_elems + alpha - 1) / alpha;
  int rc = db_setup(ctx, prm, n_elems, elem_bytes, alpha, d, dims, P0, real_units, out);
  if (rc) return rc;
  ByteSrc src{elems_dev, unit_bytes, unit_bytes, n_elems * elem_bytes};
  const uint64_t r = n_elems % alpha;
  if (r == 0 || ctx->alpha_zero_pad) {
    rc = launch_import_bytes(ctx, prm, src, real_units, (*out)->units, P0, (*out)->words);
  } else {
    // units before the last as they lie; the last one staged as the reference fills it
    rc = launch_import_bytes(ctx, prm, src, real_units - 1, real_units - 1, P0, (*out)->words);
    if (!rc) rc = import_partial_aggregate(ctx, prm, elems_dev, n_elems, elem_bytes, alpha, P0, *out);
  }
  if (!rc) rc = db_to_matrix(ctx, *out);
  if (rc) {
    xpg_db_free(*out);
    *out = nullptr;
  }
  return rc;
}

int xpg_db_import(xpg_ctx *ctx, const xpg_params *prm, const uint8_t *elems, uint64_t n_elems,
                  uint64_t elem_bytes, uint32_t alpha, uint32_t d, const uint32_t *dims, xpg_db **out) {
  if (!ctx || !elems) return XPG_EINVAL;
  const uint64_t bytes = n_elems * elem_bytes;
  uint8_t *tmp = nullptr;
  (void)hipSetDevice(ctx->device);
  HIP_OK(hipMalloc((void **)&tmp, bytes + 16), "staging allocation");
  int rc = copy_h2d_staged(ctx, tmp, elems, bytes);
  if (!rc) rc = xpg_db_import_dev(ctx, prm, tmp, n_elems, elem_bytes, alpha, d, dims, out);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(tmp);
  return rc;
}

int xpg_db_import_ntt_words(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *words, uint64_t units,
                            uint32_t P, uint64_t raw_elem_bytes, uint32_t d, const uint32_t *dims, xpg_db **out) {
  if (!ctx || !words) return XPG_EINVAL;
  int rc = db_setup(ctx, prm, units, raw_elem_bytes, 1, d, dims, P, units, out);
  if (rc) return rc;
  xpg_db *db = *out;
  const uint64_t per = (uint64_t)P * prm->m * prm->n;
  rc = copy_h2d_staged(ctx, db->words, words, units * per * 8);
  if (!rc) rc = launch_fill_zero(ctx, db->words + units * per, (db->units - units) * per);
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = set_err(ctx, XPG_EDEVICE, "sync");
  if (!rc) rc = db_to_matrix(ctx, db);
  if (rc) {
    xpg_db_free(db);
    *out = nullptr;
  }
  return rc;
}

int xpg_db_export_ntt_words(xpg_ctx *ctx, const xpg_db *db, uint64_t *out, uint64_t cap_words) {
  if (!ctx || !db || !out) return XPG_EINVAL;
  if (!db->words) return set_err(ctx, XPG_EINVAL, "a database in the matrix layout has no row-major words");
  const uint64_t words = db->units * db->P0 * (uint64_t)db->prm.m * db->prm.n;
  if (cap_words < words) return set_err(ctx, XPG_ESIZE, "export buffer too small");
  HIP_OK(hipMemcpyAsync(out, db->words, words * 8, hipMemcpyDeviceToHost, ctx->stream), "D2H copy");
  HIP_OK(hipStreamSynchronize(ctx->stream), "stream sync");
  return XPG_OK;
}

// ---- the NTT database in a file (no reference counterpart: the reference imports at
// every start and keeps the imported database in RAM, apps/server/PIRServer.cpp:71-104) ----
//
// Layout: a 512-byte header (host byte order), then units x P0 x m x n words exactly as
// they lie in HBM.  The header holds the parameters, the grid and sizes, the
// partial-aggregate convention, the caller's 32-byte digest of the raw database, and an
// order-keyed checksum of the words (launch_word_checksum) that the load recomputes on
// the device.
namespace {
constexpr char kNttMagic[8] = {'X', 'P', 'G', 'N', 'T', 'T', 'D', 'B'};
constexpr uint32_t kNttVersion = 1;
constexpr size_t kNttHeader = 512;

struct NttFileHeader {
  char magic[8];
  uint32_t version, header_bytes;
  uint32_t k, n, logn, m, bits, Berr;
  int32_t A;
  uint32_t alpha;
  uint64_t n0, moduli[4], abs_bytes, ct_bytes;
  uint64_t n_elems, elem_bytes, units, real_units, words, checksum;
  uint32_t d, P0, alpha_zero_pad, reserved;
  uint32_t dims[XPG_MAX_LEVELS];
  uint8_t digest[32];
};
static_assert(sizeof(NttFileHeader) <= kNttHeader, "header fits its block");

bool same_params(const NttFileHeader &h, const xpg_params *p) {
  if (h.k != p->k || h.n != p->n || h.logn != p->logn || h.m != p->m || h.bits != p->bits || h.Berr != p->Berr ||
      h.A != p->A || h.n0 != p->n0 || h.abs_bytes != p->abs_bytes || h.ct_bytes != p->ct_bytes)
    return false;
  for (uint32_t c = 0; c < p->m && c < 4; c++)
    if (h.moduli[c] != p->moduli[c]) return false;
  return true;
}

// a checksum of words on the device, read back
int device_checksum(xpg_ctx *ctx, const uint64_t *words, uint64_t n, uint64_t *out) {
  uint64_t *dev = nullptr;
  HIP_OK(hipMalloc((void **)&dev, 256), "checksum word");
  int rc = launch_word_checksum(ctx, words, n, dev);
  if (!rc && (hipMemcpyAsync(out, dev, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
              hipStreamSynchronize(ctx->stream) != hipSuccess))
    rc = set_err(ctx, XPG_EDEVICE, "checksum read-back");
  (void)hipFree(dev);
  return rc;
}

struct Fd {
  int fd = -1;
  ~Fd() {
    if (fd >= 0) close(fd);
  }
};

bool write_all(int fd, const void *buf, size_t len) {
  const char *p = static_cast<const char *>(buf);
  while (len) {
    const ssize_t w = write(fd, p, len);
    if (w < 0 && errno == EINTR) continue;
    if (w <= 0) return false;
    p += w;
    len -= (size_t)w;
  }
  return true;
}

// bytes read (short only at the end of the file), or -1 on an error
ssize_t read_full(int fd, void *buf, size_t len, off_t at) {
  char *p = static_cast<char *>(buf);
  size_t got = 0;
  while (got < len) {
    const ssize_t r = pread(fd, p + got, len - got, at + (off_t)got);
    if (r < 0 && errno == EINTR) continue;
    if (r < 0) return -1;
    if (r == 0) break;
    got += (size_t)r;
  }
  return (ssize_t)got;
}

int ensure_pin(xpg_ctx *ctx) {
  for (int b = 0; b < 2; b++) {
    if (!ctx->pin[b]) HIP_OK(hipHostMalloc(&ctx->pin[b], kPinChunk, hipHostMallocDefault), "pinned staging chunk");
    if (!ctx->pin_ev[b]) HIP_OK(hipEventCreateWithFlags(&ctx->pin_ev[b], hipEventDisableTiming), "staging event");
  }
  return XPG_OK;
}
}  // namespace

int xpg_db_save(xpg_ctx *ctx, const xpg_db *db, const char *path, const uint8_t *digest) {
  if (!ctx || !db || !path) return XPG_EINVAL;
  if (!db->words) return set_err(ctx, XPG_EINVAL, "a database in the matrix layout has no row-major words to save");
  if (db->device != ctx->device) return set_err(ctx, XPG_EINVAL, "database lives on another GPU");
  NttFileHeader h{};
  memcpy(h.magic, kNttMagic, 8);
  h.version = kNttVersion;
  h.header_bytes = (uint32_t)kNttHeader;
  const xpg_params &p = db->prm;
  h.k = p.k, h.n = p.n, h.logn = p.logn, h.m = p.m, h.bits = p.bits, h.Berr = p.Berr, h.A = p.A;
  h.n0 = p.n0, h.abs_bytes = p.abs_bytes, h.ct_bytes = p.ct_bytes;
  for (int c = 0; c < 4; c++) h.moduli[c] = p.moduli[c];
  h.alpha = db->alpha;
  h.n_elems = db->n_elems, h.elem_bytes = db->elem_bytes, h.units = db->units, h.real_units = db->real_units;
  h.words = db->units * db->P0 * (uint64_t)p.m * p.n;
  h.d = db->d, h.P0 = db->P0, h.alpha_zero_pad = (uint32_t)db->alpha_zero_pad;
  for (uint32_t l = 0; l < db->d; l++) h.dims[l] = db->dims[l];
  if (digest) memcpy(h.digest, digest, 32);
  int rc = device_checksum(ctx, db->words, h.words, &h.checksum);
  if (rc) return rc;
  rc = ensure_pin(ctx);
  if (rc) return rc;
  // written under a temporary name of its own (process, call) and renamed once complete:
  // a reader never sees a partial file under `path`, and two writers never share one
  static std::atomic<uint64_t> saves{0};
  const std::string tmp =
      std::string(path) + ".tmp." + std::to_string((long long)getpid()) + "." + std::to_string(saves++);
  Fd f;
  f.fd = open(tmp.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (f.fd < 0) return set_err(ctx, XPG_EINVAL, std::string("cannot create ") + tmp + ": " + strerror(errno));
  char block[kNttHeader] = {0};
  memcpy(block, &h, sizeof(h));
  bool ok = write_all(f.fd, block, kNttHeader);
  // device -> two pinned chunks -> file: chunk i+1's copy runs while chunk i is written
  const uint64_t bytes = h.words * 8;
  const size_t nch = (size_t)((bytes + kPinChunk - 1) / kPinChunk);
  auto enqueue = [&](size_t i) {
    const size_t off = i * kPinChunk, len = std::min<uint64_t>(kPinChunk, bytes - off);
    const int b = (int)(i & 1);
    return hipMemcpyAsync(ctx->pin[b], reinterpret_cast<const char *>(db->words) + off, len, hipMemcpyDeviceToHost,
                          ctx->stream) == hipSuccess &&
           hipEventRecord(ctx->pin_ev[b], ctx->stream) == hipSuccess;
  };
  if (ok && nch) ok = enqueue(0);
  for (size_t i = 0; ok && i < nch; i++) {
    if (i + 1 < nch) ok = enqueue(i + 1);
    const int b = (int)(i & 1);
    const size_t off = i * kPinChunk, len = std::min<uint64_t>(kPinChunk, bytes - off);
    ok = ok && hipEventSynchronize(ctx->pin_ev[b]) == hipSuccess && write_all(f.fd, ctx->pin[b], len);
    // the buffer is free once written; the copy two chunks on reuses it
  }
  (void)hipStreamSynchronize(ctx->stream);
  ok = ok && fsync(f.fd) == 0;
  close(f.fd);
  f.fd = -1;
  if (!ok || rename(tmp.c_str(), path) != 0) {
    (void)unlink(tmp.c_str());
    return set_err(ctx, XPG_EDEVICE, std::string("writing ") + path + " failed");
  }
  return XPG_OK;
}

int xpg_db_load(xpg_ctx *ctx, const xpg_params *prm, const char *path, const uint8_t *digest, xpg_db **out) {
  if (!ctx || !prm || !path || !out) return XPG_EINVAL;
  *out = nullptr;
  Fd f;
  f.fd = open(path, O_RDONLY);
  if (f.fd < 0) return set_err(ctx, XPG_EINVAL, std::string("cannot open ") + path + ": " + strerror(errno));
  char block[kNttHeader];
  if (read_full(f.fd, block, kNttHeader, 0) != (ssize_t)kNttHeader)
    return set_err(ctx, XPG_ESIZE, "database file shorter than its header");
  NttFileHeader h;
  memcpy(&h, block, sizeof(h));
  if (memcmp(h.magic, kNttMagic, 8) != 0 || h.version != kNttVersion || h.header_bytes != kNttHeader)
    return set_err(ctx, XPG_EINVAL, "not a database file of this version");
  if (!same_params(h, prm)) return set_err(ctx, XPG_EINVAL, "database file made under other parameters");
  if (digest && memcmp(h.digest, digest, 32) != 0)
    return set_err(ctx, XPG_EINVAL, "database file made from another database (digest differs)");
  if (h.d == 0 || h.d > XPG_MAX_LEVELS || h.alpha == 0 || h.P0 == 0)
    return set_err(ctx, XPG_EINVAL, "database file header out of range");
  if (h.words != prod_dims(h.d, h.dims, 0) * h.P0 * (uint64_t)prm->m * prm->n)
    return set_err(ctx, XPG_EINVAL, "database file header sizes disagree");
  struct stat st;
  if (fstat(f.fd, &st) != 0 || (uint64_t)st.st_size != kNttHeader + h.words * 8)
    return set_err(ctx, XPG_ESIZE, "database file size differs from its header's");
  xpg_db *db = nullptr;
  int rc = db_setup(ctx, prm, h.n_elems, h.elem_bytes, h.alpha, h.d, h.dims, h.P0, h.real_units, &db);
  if (rc) return rc;
  db->alpha_zero_pad = (int)h.alpha_zero_pad;
  rc = ensure_pin(ctx);
  // file -> two pinned chunks -> device: chunk i is read while chunk i-1's copy runs
  const uint64_t bytes = h.words * 8;
  for (uint64_t off = 0, i = 0; !rc && off < bytes; off += kPinChunk, i++) {
    const int b = (int)(i & 1);
    const size_t len = std::min<uint64_t>(kPinChunk, bytes - off);
    if (hipEventSynchronize(ctx->pin_ev[b]) != hipSuccess) rc = set_err(ctx, XPG_EDEVICE, "staging wait");
    if (!rc && read_full(f.fd, ctx->pin[b], len, (off_t)(kNttHeader + off)) != (ssize_t)len)
      rc = set_err(ctx, XPG_ESIZE, "database file read failed");
    if (!rc && (hipMemcpyAsync(reinterpret_cast<char *>(db->words) + off, ctx->pin[b], len, hipMemcpyHostToDevice,
                               ctx->stream) != hipSuccess ||
                hipEventRecord(ctx->pin_ev[b], ctx->stream) != hipSuccess))
      rc = set_err(ctx, XPG_EDEVICE, "H2D copy");
  }
  uint64_t sum = 0;
  if (!rc) rc = device_checksum(ctx, db->words, h.words, &sum);
  if (!rc && sum != h.checksum) rc = set_err(ctx, XPG_EINVAL, "database file checksum mismatch (corrupted words)");
  if (!rc) rc = db_to_matrix(ctx, db);
  if (rc) {
    (void)hipStreamSynchronize(ctx->stream);
    xpg_db_free(db);
    return rc;
  }
  *out = db;
  return XPG_OK;
}

uint32_t xpg_db_polys(const xpg_db *db) { return db ? db->P0 : 0; }
uint32_t xpg_db_reply_polys(const xpg_db *db) { return db ? db->Pl[db->d] : 0; }
uint32_t xpg_db_levels(const xpg_db *db, uint32_t *dims, uint32_t cap) {
  if (!db) return 0;
  for (uint32_t l = 0; dims && l < db->d && l < cap; l++) dims[l] = db->dims[l];
  return db->d;
}
uint64_t xpg_db_raw_bytes(const xpg_db *db) { return db ? db->n_elems * db->elem_bytes : 0; }
uint64_t xpg_db_query_words(const xpg_db *db) {
  if (!db) return 0;
  uint64_t s = 0;
  for (uint32_t l = 0; l < db->d; l++) s += db->dims[l];
  return s * 2 * db->prm.m * db->prm.n;
}

void xpg_db_free(xpg_db *db) {
  if (!db) return;
  (void)hipSetDevice(db->device);
  if (db->ctx) {
    (void)hipStreamSynchronize(db->ctx->stream);
    graph_cache_clear(db->ctx, db->serial);  // other contexts' entries never match again
    db->ctx->dbs.erase(db);
  }
  if (db->words) (void)hipFree(db->words);
  if (db->mx) (void)hipFree(db->mx);
  if (db->mx_tail) (void)hipFree(db->mx_tail);
  delete db;
}

// ---- replies (PIRReplyGeneratorNFL_internal.cpp:340-445) ----
// Bytes of the two intermediate-level buffers of levels [from, d): the largest level
// reply (groups x P x 2 x m x n) and the largest next-level NTT database.
static void level_buf_sizes(const xpg_db *db, uint32_t from, size_t *need_reply, size_t *need_db) {
  const uint64_t mn = (uint64_t)db->prm.m * db->prm.n;
  *need_reply = *need_db = 0;
  for (uint32_t l = from; l + 1 < db->d; l++) {
    const uint64_t groups = prod_dims(db->d, db->dims, l + 1);
    *need_reply = std::max<size_t>(*need_reply, groups * db->Pl[l] * 2 * mn * 8);
    *need_db = std::max<size_t>(*need_db, groups * db->Pl[l + 1] * mn * 8);
  }
}

// Levels [first, d) of one query.  `data` holds level `first`'s NTT database (P polys per
// unit); `q` points at level `first`'s query part.  When first > 0 the caller passes the
// previous level's reply in `prev_reply` and this routine starts with the recursion.
static int run_levels_in(xpg_ctx *ctx, const xpg_db *db, uint32_t first, const uint64_t *data,
                         const uint64_t *q, const uint64_t *prev_reply, uint64_t *reply_dev);

// With a workspace set, the carve of one reply is returned at its end: the next reply's
// launches queue behind this one's on the same stream, so they may reuse the bytes.
static int run_levels(xpg_ctx *ctx, const xpg_db *db, uint32_t first, const uint64_t *data, const uint64_t *q,
                      const uint64_t *prev_reply, uint64_t *reply_dev) {
  const size_t used = ctx->ws_used;
  const int rc = run_levels_in(ctx, db, first, data, q, prev_reply, reply_dev);
  if (ctx->ws) ctx->ws_used = used;
  return rc;
}

static int run_levels_in(xpg_ctx *ctx, const xpg_db *db, uint32_t first, const uint64_t *data,
                         const uint64_t *q, const uint64_t *prev_reply, uint64_t *reply_dev) {
  const xpg_params *prm = &db->prm;
  const uint64_t mn = (uint64_t)prm->m * prm->n;
  // intermediate levels: lvl_reply holds a level's reply, lvl_db the next level's NTT
  // database (sized before any launch reads them): the context's level buffers, or
  // inside xpg_reply_dev_noalloc the front of its scratch array
  size_t need_reply = 0, need_db = 0;
  level_buf_sizes(db, first ? first - 1 : 0, &need_reply, &need_db);
  uint64_t *lvl_reply = nullptr, *lvl_db = nullptr;
  if (need_db && ctx->ws) {
    const size_t r_al = (need_reply + 255) & ~(size_t)255, d_al = (need_db + 255) & ~(size_t)255;
    if (r_al + d_al > ctx->ws_cap - ctx->ws_used)
      return set_err(ctx, XPG_ESIZE, "the workspace is too small for the level buffers");
    lvl_reply = reinterpret_cast<uint64_t *>(ctx->ws + ctx->ws_used);
    lvl_db = reinterpret_cast<uint64_t *>(ctx->ws + ctx->ws_used + r_al);
    ctx->ws_used += r_al + d_al;
  } else if (need_db) {
    int rc = ensure_level_buf(ctx, 0, need_reply);
    if (!rc) rc = ensure_level_buf(ctx, 1, need_db);
    if (rc) return rc;
    lvl_reply = ctx->level_buf[0];
    lvl_db = ctx->level_buf[1];
  }
  if (first > 0) {
    // fromResulttoInData on the previous level's reply
    const uint32_t l = first - 1;
    const uint64_t groups = prod_dims(db->d, db->dims, first);
    const uint64_t unit = (uint64_t)db->Pl[l] * 2 * mn * 8;
    ByteSrc src{(const uint8_t *)prev_reply, unit, unit, groups * unit};
    int rc = launch_import_bytes(ctx, prm, src, groups, groups, db->Pl[first], lvl_db);
    if (rc) return rc;
    data = lvl_db;
  }
  for (uint32_t l = first; l < db->d; l++) {
    const uint64_t groups = prod_dims(db->d, db->dims, l + 1);
    const uint32_t P = db->Pl[l];
    const bool last = (l + 1 == db->d);
    uint64_t *out = last ? reply_dev : lvl_reply;
    int rc = XPG_OK;
    if (l == 0 && db->mx) {
      void *work = nullptr;
      rc = get_scratch(ctx, mx_work_bytes(db->mxg, 1), &work);
      if (!rc) rc = launch_mx_level(ctx, prm, db->mxg, db->mx, db->mx_tail, &q, &out, 1, work);
    } else {
      rc = launch_mac_level(ctx, prm, data, q, db->dims[l], groups, P, out);
    }
    if (rc) return rc;
    q += (uint64_t)db->dims[l] * 2 * mn;
    if (!last) {
      // fromResulttoInData: each group's P ciphertexts, as bytes, become one unit
      const uint64_t unit = (uint64_t)P * 2 * mn * 8;
      ByteSrc src{(const uint8_t *)lvl_reply, unit, unit, groups * unit};
      rc = launch_import_bytes(ctx, prm, src, groups, groups, db->Pl[l + 1], lvl_db);
      if (rc) return rc;
      data = lvl_db;
    }
  }
  return XPG_OK;
}

// A reply replayed from a captured graph.  The first call for a (database, query,
// reply) triple runs eagerly, which also sizes every context buffer the levels use; the
// second captures the same launches (nothing allocates then) and replays them; later
// calls replay.  Entries die with the epoch: a moved buffer or a changed option.
static constexpr size_t kGraphCacheMax = 32;

static int reply_graph(xpg_ctx *ctx, const xpg_db *db, const uint64_t *q, uint64_t *r) {
  auto &cache = ctx->graph_cache;
  GraphEntry *hit = nullptr;
  for (auto &g : cache)
    if (g.db_serial == db->serial && g.q == q && g.r == r) hit = &g;
  ctx->use_clock++;
  const uint64_t tg = tune_generation();
  const bool current = hit && hit->epoch == ctx->epoch && hit->tune_gen == tg;
  if (current && hit->exec) {
    hit->last_use = ctx->use_clock;
    ctx->graph_launches++;
    HIP_OK(hipGraphLaunch(hit->exec, ctx->stream), "graph launch");
    return XPG_OK;
  }
  if (!current) {
    // first sight (or stale): run eagerly, remember the triple
    const int rc = run_levels(ctx, db, 0, db->words, q, nullptr, r);
    if (rc) return rc;
    if (!hit) {
      if (cache.size() >= kGraphCacheMax) {
        size_t lru = 0;
        for (size_t i = 1; i < cache.size(); i++)
          if (cache[i].last_use < cache[lru].last_use) lru = i;
        if (cache[lru].exec) (void)hipGraphExecDestroy(cache[lru].exec);
        cache[lru] = cache.back();
        cache.pop_back();
      }
      cache.push_back(GraphEntry{});
      hit = &cache.back();
    } else if (hit->exec) {
      (void)hipGraphExecDestroy(hit->exec);
    }
    *hit = GraphEntry{db->serial, q, r, ctx->epoch, tg, nullptr, ctx->use_clock};
    return XPG_OK;
  }
  // second sight at this epoch: capture
  const uint64_t epoch = ctx->epoch;
  HIP_OK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  const int rc = run_levels(ctx, db, 0, db->words, q, nullptr, r);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
  hipGraphExec_t exec = nullptr;
  if (!rc && ec == hipSuccess && graph && ctx->epoch == epoch && tune_generation() == tg &&
      hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
    hit->exec = exec;
    hit->last_use = ctx->use_clock;
    ctx->graph_captures++;
  }
  if (graph) (void)hipGraphDestroy(graph);
  if (!hit->exec) {
    // capture failed (nothing ran): run eagerly, which also reports any real error
    hit->epoch = 0;
    (void)hipGetLastError();
    return run_levels(ctx, db, 0, db->words, q, nullptr, r);
  }
  ctx->graph_launches++;
  HIP_OK(hipGraphLaunch(hit->exec, ctx->stream), "graph launch");
  return XPG_OK;
}

int xpg_reply_dev(xpg_ctx *ctx, const xpg_db *db, const uint64_t *query_dev, uint64_t *reply_dev) {
  if (!ctx || !db || !query_dev || !reply_dev) return set_err(ctx, XPG_EINVAL, "null argument");
  if (db->device != ctx->device) return set_err(ctx, XPG_EINVAL, "database lives on another GPU");
  if (ctx->graphs && !ctx->timing) return reply_graph(ctx, db, query_dev, reply_dev);
  return run_levels(ctx, db, 0, db->words, query_dev, nullptr, reply_dev);
}

// The reply of xpg_reply_dev without any allocation: the levels run eagerly on the
// context stream, their intermediate buffers and block summaries carved from the
// device's static scratch array (level buffers first, each launch's summaries after
// them).  Every need is checked before the first launch, so a reply that does not fit
// fails with nothing enqueued.
int xpg_reply_dev_noalloc(xpg_ctx *ctx, const xpg_db *db, const uint64_t *query_dev, uint64_t *reply_dev) {
  if (!ctx || !db || !query_dev || !reply_dev) return set_err(ctx, XPG_EINVAL, "null argument");
  if (db->device != ctx->device) return set_err(ctx, XPG_EINVAL, "database lives on another GPU");
  hipError_t herr = hipSuccess;
  uint8_t *base = scratch_of(ctx->device, &herr);
  if (!base)
    return set_err(ctx, XPG_EDEVICE, std::string("no scratch array address on this device: ") + hipGetErrorString(herr));
  const xpg_params *prm = &db->prm;
  size_t need_reply = 0, need_db = 0;
  level_buf_sizes(db, 0, &need_reply, &need_db);
  const size_t lvl = need_db ? ((need_reply + 255) & ~(size_t)255) + ((need_db + 255) & ~(size_t)255) : 0;
  size_t scr = 0;
  for (uint32_t l = 0; l < db->d; l++) {
    const uint64_t groups = prod_dims(db->d, db->dims, l + 1);
    if (l == 0 && db->mx) {
      scr = std::max<size_t>(scr, mx_work_bytes(db->mxg, 1));
      continue;
    }
    const LevelCfg cfg = resolve_level_cfg(ctx, prm, db->dims[l], groups, db->Pl[l]);
    scr = std::max<size_t>(scr, level_scratch_bytes(ctx, prm, db->dims[l], groups, db->Pl[l], cfg));
  }
  if (lvl > XPG_NOALLOC_SCRATCH_BYTES || scr > XPG_NOALLOC_SCRATCH_BYTES - lvl)
    return set_err(ctx, XPG_ESIZE,
                   "reply needs " + std::to_string(lvl + scr) + " bytes of scratch; the no-allocation entry point has " +
                       std::to_string((uint64_t)XPG_NOALLOC_SCRATCH_BYTES));
  if (db->d > 1 && !ctx->tables.count(((uint64_t)prm->n << 8) | prm->m))
    return set_err(ctx, XPG_EINVAL,
                   "no NTT tables for these parameters on this context: the recursion needs them (an import or "
                   "an ordinary reply on this context builds them)");
  (void)hipSetDevice(ctx->device);
  ctx->ws = base;
  ctx->ws_cap = XPG_NOALLOC_SCRATCH_BYTES;
  ctx->ws_used = 0;
  ctx->noalloc = true;
  const int rc = run_levels(ctx, db, 0, db->words, query_dev, nullptr, reply_dev);
  ctx->ws = nullptr;
  ctx->ws_cap = ctx->ws_used = 0;
  ctx->noalloc = false;
  return rc;
}

// Level 0 of every request on one database in one batched launch, then each request's
// remaining levels.  Worth it when level 0 has enough chains to fill the GPU per query
// group (the batched kernel runs one chain per lane).
static int batch_chunk(xpg_ctx *ctx, const xpg_db *db, const std::vector<const xpg_request *> &reqs);
// device bytes the grouped levels of one batch may take (level databases + outputs);
// option "group_budget" overrides it
static const uint64_t kGroupedBudget = 24ull << 30;

// Whether a batch of qn queries on the database's level 0 runs as one batched launch:
// the tuning table's measurement when it has one for this shape, else a fixed rule
// (enough lanes per query group to fill the GPU).
static bool batch_pays(const xpg_ctx *ctx, const xpg_db *db, uint32_t qn) {
  if (qn < 2) return false;
  const xpg_params *prm = &db->prm;
  const uint64_t groups0 = prod_dims(db->d, db->dims, 1);
  LevelCfg c;
  const TuneKey key{1, prm->n, prm->m, db->P0, db->dims[0], groups0, tune_qn_bucket(qn)};
  if (!options_pinned(ctx) && tune_lookup(ctx->device, key, &c)) return c.kernel != 0;
  const uint64_t lanes = groups0 * db->P0 * (uint64_t)prm->m * prm->n;
  return lanes * ((qn + 7) / 8) >= (1ull << 16);
}

// A workspace carve is returned when the scope that made it ends (later calls queue
// behind this one on the stream, so they may reuse the bytes).
struct WsMark {
  xpg_ctx *c;
  size_t used;
  ~WsMark() {
    if (c->ws) c->ws_used = used;
  }
};

// Bytes of one buffer of a batched call: carved from the call's workspace when one is set
// (kept until the call's end), else the context's grown buffer (which: 0 = level reply,
// 1 = level database, 2 = level-0 outputs).
static int batch_buffer(xpg_ctx *ctx, int which, size_t bytes, uint64_t **p) {
  if (ctx->ws) {
    const size_t al = (bytes + 255) & ~(size_t)255;
    if (al > ctx->ws_cap - ctx->ws_used)
      return set_err(ctx, XPG_ESIZE, "the workspace is too small for the batch's intermediate levels");
    *p = reinterpret_cast<uint64_t *>(ctx->ws + ctx->ws_used);
    ctx->ws_used += al;
    return XPG_OK;
  }
  const int rc = which == 2 ? ensure_batch_buf(ctx, bytes) : ensure_level_buf(ctx, which, bytes);
  if (rc) return rc;
  *p = which == 2 ? ctx->batch_buf : ctx->level_buf[which];
  return XPG_OK;
}

// bytes of the buffers one request adds to a batch's levels 1..d-1 (the level database
// and, for d > 2, a level's outputs)
static void batch_level_bytes(const xpg_db *db, size_t *per_db, size_t *per_out) {
  const uint64_t mn = (uint64_t)db->prm.m * db->prm.n;
  *per_db = *per_out = 0;
  for (uint32_t l = 1; l < db->d; l++) {
    *per_db = std::max<size_t>(*per_db, prod_dims(db->d, db->dims, l) * db->Pl[l] * mn * 8);
    if (l + 1 < db->d)
      *per_out = std::max<size_t>(*per_out, prod_dims(db->d, db->dims, l + 1) * db->Pl[l] * 2 * mn * 8);
  }
}

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

extern "C++" uint64_t batch_workspace_bytes(const xpg_ctx *ctx, const xpg_db *db, uint32_t k) {
  const xpg_params *prm = &db->prm;
  const uint64_t mn = (uint64_t)prm->m * prm->n;
  // batched: every request's level-0 outputs, then one request's later levels at least
  const uint64_t lvl0 = db->d > 1 ? al256(k * prod_dims(db->d, db->dims, 1) * db->P0 * 2 * mn * 8) : 0;
  size_t per_db = 0, per_out = 0;
  batch_level_bytes(db, &per_db, &per_out);
  // (plus the fused launch's tile table)
  // (plus, in the matrix layout, the launch's query fragments and staged residues)
  const uint64_t batched = lvl0 + al256(fused_table_bytes(k)) +
                           std::max<uint64_t>(db->mx ? al256(mx_work_bytes(db->mxg, k)) : 0,
                                              db->d > 1 ? al256(per_db) + (per_out ? al256(per_out) : 0) : 0);
  // one reply at a time: its level buffers and the largest level's scratch
  size_t need_reply = 0, need_db = 0;
  level_buf_sizes(db, 0, &need_reply, &need_db);
  uint64_t single = need_db ? al256(need_reply) + al256(need_db) : 0, scr = 0;
  for (uint32_t l = 0; l < db->d; l++) {
    const uint64_t groups = prod_dims(db->d, db->dims, l + 1);
    if (l == 0 && db->mx) {
      scr = std::max<uint64_t>(scr, mx_work_bytes(db->mxg, 1));
      continue;
    }
    const LevelCfg cfg = resolve_level_cfg(ctx, prm, db->dims[l], groups, db->Pl[l]);
    scr = std::max<uint64_t>(scr, level_scratch_bytes(ctx, prm, db->dims[l], groups, db->Pl[l], cfg));
  }
  return std::max<uint64_t>(batched, single + scr);
}

// Levels 1..d-1 of a chunk's requests, whose level-0 outputs lie side by side at lvl0.
static int batch_levels(xpg_ctx *ctx, const xpg_db *db, const std::vector<const xpg_request *> &reqs,
                        const uint64_t *lvl0);

// One batched launch for level 0 of up to 64 requests, then each request's other levels.
static int batch_chunk(xpg_ctx *ctx, const xpg_db *db, const std::vector<const xpg_request *> &reqs) {
  const xpg_params *prm = &db->prm;
  const uint64_t mn = (uint64_t)prm->m * prm->n;
  const uint64_t groups0 = prod_dims(db->d, db->dims, 1);
  const uint32_t qn = (uint32_t)reqs.size();
  const uint64_t lvl0_words = groups0 * db->P0 * 2 * mn;
  // a workspace carve is returned at the call's end (later calls queue behind this one)
  WsMark mark{ctx, ctx->ws_used};
  // level-0 outputs: the final replies when d == 1, else per-request intermediates
  uint64_t *lvl0 = nullptr;
  if (db->d > 1) {
    const int rc0 = batch_buffer(ctx, 2, qn * lvl0_words * 8, &lvl0);
    if (rc0) return rc0;
  }
  std::vector<const uint64_t *> qp(qn);
  std::vector<uint64_t *> op(qn);
  for (uint32_t i = 0; i < qn; i++) {
    qp[i] = reqs[i]->query_dev;
    op[i] = db->d == 1 ? reqs[i]->reply_dev : lvl0 + i * lvl0_words;
  }
  int rc = XPG_OK;
  if (db->mx) {
    // the launch's workspace after the level-0 outputs (not kept: the levels below reuse it)
    void *work = nullptr;
    rc = get_scratch(ctx, mx_work_bytes(db->mxg, qn), &work);
    if (!rc) rc = launch_mx_level(ctx, prm, db->mxg, db->mx, db->mx_tail, qp.data(), op.data(), qn, work);
  } else {
    rc = launch_mac_batch(ctx, prm, db->words, qp.data(), op.data(), qn, db->dims[0], groups0, db->P0);
  }
  if (rc || db->d == 1) return rc;
  return batch_levels(ctx, db, reqs, lvl0);
}

static int batch_levels(xpg_ctx *ctx, const xpg_db *db, const std::vector<const xpg_request *> &reqs,
                        const uint64_t *lvl0) {
  // Levels 1..d-1 of all the chunk's requests together: per level, one recursion import
  // over every request's previous-level reply (they lie side by side) and one MAC launch
  // over the requests' level databases side by side, each request reading its own query.
  // Requests are taken in sub-chunks whose level buffers fit the budget: the rest of the
  // workspace when one is set, else kGroupedBudget (option group_budget).
  const xpg_params *prm = &db->prm;
  const uint64_t mn = (uint64_t)prm->m * prm->n;
  const uint32_t qn = (uint32_t)reqs.size(), d = db->d;
  const uint64_t lvl0_words = prod_dims(d, db->dims, 1) * db->P0 * 2 * mn;
  WsMark mark{ctx, ctx->ws_used};
  size_t per_db = 0, per_out = 0;
  batch_level_bytes(db, &per_db, &per_out);
  const size_t per_req = al256(per_db) + (per_out ? al256(per_out) : 0);
  const uint64_t budget = ctx->ws ? ctx->ws_cap - ctx->ws_used : ctx->group_budget ? ctx->group_budget : kGroupedBudget;
  const uint32_t sub = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(qn, budget / per_req));
  uint64_t *lvl_db = nullptr, *lvl_out = nullptr;
  int rc = XPG_OK;
  if (ctx->ws) {  // carved once, for the largest sub-chunk
    rc = batch_buffer(ctx, 1, sub * per_db, &lvl_db);
    if (!rc && per_out) rc = batch_buffer(ctx, 0, sub * per_out, &lvl_out);
    if (rc) return rc;
  }
  for (uint32_t s0 = 0; s0 < qn; s0 += sub) {
    const uint32_t cnt = std::min<uint32_t>(sub, qn - s0);
    if (!ctx->ws) {
      rc = batch_buffer(ctx, 1, cnt * per_db, &lvl_db);
      if (!rc && per_out) rc = batch_buffer(ctx, 0, cnt * per_out, &lvl_out);
      if (rc) return rc;
    }
    const uint64_t *prev = lvl0 + s0 * lvl0_words;
    uint64_t qoff = (uint64_t)db->dims[0] * 2 * mn;
    std::vector<const uint64_t *> qs(cnt);
    std::vector<uint64_t *> outs(cnt);
    for (uint32_t l = 1; l < d; l++) {
      const uint64_t groups_prev = prod_dims(d, db->dims, l), groups_l = prod_dims(d, db->dims, l + 1);
      const uint64_t unit = (uint64_t)db->Pl[l - 1] * 2 * mn * 8;
      const uint64_t units = cnt * groups_prev;
      ByteSrc src{(const uint8_t *)prev, unit, unit, units * unit};
      rc = launch_import_bytes(ctx, prm, src, units, units, db->Pl[l], lvl_db);
      if (rc) return rc;
      for (uint32_t i = 0; i < cnt; i++) {
        qs[i] = reqs[s0 + i]->query_dev + qoff;
        outs[i] = reqs[s0 + i]->reply_dev;
      }
      const bool last = l + 1 == d;
      rc = launch_mac_level_multi(ctx, prm, lvl_db, qs.data(), cnt, db->dims[l], groups_l, db->Pl[l],
                                  last ? outs.data() : nullptr, last ? nullptr : lvl_out);
      if (rc) return rc;
      prev = lvl_out;
      qoff += (uint64_t)db->dims[l] * 2 * mn;
    }
  }
  return XPG_OK;
}

// Chunks (up to 64 requests of one database each) whose level 0 runs register-blocked,
// in fused launches: as many chunks per launch as the launch's segment table and the
// level-0 budget hold (the rest of the workspace when one is set, else kGroupedBudget),
// then each chunk's remaining levels.
struct BatchChunk {
  const xpg_db *db;
  std::vector<const xpg_request *> reqs;
};

static uint64_t lvl0_bytes(const xpg_db *db, uint32_t qn) {
  return db->d > 1 ? al256((uint64_t)qn * prod_dims(db->d, db->dims, 1) * db->P0 * 2 * db->prm.m * db->prm.n * 8) : 0;
}

static int fused_table(xpg_ctx *ctx, size_t bytes, void **p) {
  if (ctx->ws) {
    uint64_t *w = nullptr;
    const int rc = batch_buffer(ctx, 0, bytes, &w);  // carved (`which` only names grown buffers)
    *p = w;
    return rc;
  }
  if (ctx->fused_bytes < bytes) {
    if (ctx->fused_buf) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(ctx->fused_buf);
    }
    ctx->fused_buf = nullptr;
    ctx->fused_bytes = 0;
    ctx->epoch++;
    const size_t want = std::max<size_t>(bytes, 64 << 10);
    if (hipMalloc(&ctx->fused_buf, want) != hipSuccess) return set_err(ctx, XPG_ENOMEM, "fused tile table");
    ctx->fused_bytes = want;
  }
  *p = ctx->fused_buf;
  return XPG_OK;
}

static int batch_fused(xpg_ctx *ctx, const std::vector<BatchChunk> &chunks) {
  for (size_t i = 0; i < chunks.size();) {
    WsMark mark{ctx, ctx->ws_used};
    const uint64_t avail =
        ctx->ws ? ctx->ws_cap - ctx->ws_used : ctx->group_budget ? ctx->group_budget : kGroupedBudget;
    uint32_t segs = 0, queries = 0;
    uint64_t lvl0 = 0, levels = 0;
    size_t j = i;
    for (; j < chunks.size(); j++) {
      const BatchChunk &c = chunks[j];
      const uint32_t qn = (uint32_t)c.reqs.size();
      size_t per_db = 0, per_out = 0;
      batch_level_bytes(c.db, &per_db, &per_out);
      const uint64_t lv = c.db->d > 1 ? al256(per_db) + (per_out ? al256(per_out) : 0) : 0;
      // with a workspace, the table and one request's later levels come from it too
      const uint64_t need = lvl0 + lvl0_bytes(c.db, qn) +
                            (ctx->ws ? al256(fused_table_bytes(queries + qn)) + std::max(levels, lv) : 0);
      if (j > i && (segs + fused_segments(qn) > kFusedMaxSegments || need > avail)) break;
      segs += fused_segments(qn);
      queries += qn;
      lvl0 += lvl0_bytes(c.db, qn);
      levels = std::max(levels, lv);
    }
    uint64_t *lvl0_base = nullptr;
    int rc = lvl0 ? batch_buffer(ctx, 2, lvl0, &lvl0_base) : XPG_OK;
    void *table = nullptr;
    if (!rc) rc = fused_table(ctx, fused_table_bytes(queries), &table);
    if (rc) return rc;
    std::vector<FusedPart> parts;
    std::vector<std::vector<const uint64_t *>> qp(j - i);
    std::vector<std::vector<uint64_t *>> op(j - i);
    std::vector<const uint64_t *> chunk_lvl0(j - i, nullptr);
    uint64_t off = 0;  // words into lvl0_base
    for (size_t k = i; k < j; k++) {
      const xpg_db *db = chunks[k].db;
      const uint32_t qn = (uint32_t)chunks[k].reqs.size();
      const uint64_t words = prod_dims(db->d, db->dims, 1) * db->P0 * 2 * db->prm.m * db->prm.n;
      chunk_lvl0[k - i] = lvl0_base + off;
      for (uint32_t r = 0; r < qn; r++) {
        qp[k - i].push_back(chunks[k].reqs[r]->query_dev);
        op[k - i].push_back(db->d == 1 ? chunks[k].reqs[r]->reply_dev : lvl0_base + off + r * words);
      }
      off += lvl0_bytes(db, qn) / 8;
      parts.push_back(FusedPart{&db->prm, db->words, qp[k - i].data(), op[k - i].data(), qn, db->dims[0],
                                prod_dims(db->d, db->dims, 1), db->P0});
    }
    // the part with the most moduli first (every part's moduli are its first m)
    size_t widest = 0;
    for (size_t k = 1; k < parts.size(); k++)
      if (parts[k].prm->m > parts[widest].prm->m) widest = k;
    std::swap(parts[0], parts[widest]);
    rc = launch_mac_fused(ctx, parts.data(), (uint32_t)parts.size(), table);
    if (rc) return rc;
    for (size_t k = i; k < j; k++)
      if (chunks[k].db->d > 1) {
        rc = batch_levels(ctx, chunks[k].db, chunks[k].reqs, chunk_lvl0[k - i]);
        if (rc) return rc;
      }
    i = j;
  }
  return XPG_OK;
}

extern "C++" int reply_batch_in_workspace(xpg_ctx *ctx, const xpg_request *reqs, uint32_t n_reqs, uint8_t *base,
                                         size_t bytes) {
  ctx->ws = base;
  ctx->ws_cap = bytes;
  ctx->ws_used = 0;
  const int rc = xpg_reply_batch_dev(ctx, reqs, n_reqs);
  ctx->ws = nullptr;
  ctx->ws_cap = ctx->ws_used = 0;
  return rc;
}

int xpg_reply_batch_dev(xpg_ctx *ctx, const xpg_request *reqs, uint32_t n_reqs) {
  if (!ctx || (!reqs && n_reqs)) return XPG_EINVAL;
  // group requests by database, in order of first appearance
  std::vector<const xpg_db *> order;
  std::map<const xpg_db *, std::vector<const xpg_request *>> by_db;
  for (uint32_t i = 0; i < n_reqs; i++) {
    if (!reqs[i].db || !reqs[i].query_dev || !reqs[i].reply_dev) return set_err(ctx, XPG_EINVAL, "null request field");
    if (reqs[i].db->device != ctx->device) return set_err(ctx, XPG_EINVAL, "database lives on another GPU");
    if (!by_db.count(reqs[i].db)) order.push_back(reqs[i].db);
    by_db[reqs[i].db].push_back(&reqs[i]);
  }
  // chunks of up to 64 requests per database: those whose level 0 runs register-blocked
  // go to the fused launches (option batch_fuse, on by default), the others run alone
  std::vector<BatchChunk> fused;
  for (const xpg_db *db : order) {
    const std::vector<const xpg_request *> &v = by_db[db];
    for (size_t s = 0; s < v.size(); s += 64) {
      BatchChunk c{db, std::vector<const xpg_request *>(v.begin() + s, v.begin() + std::min(v.size(), s + 64))};
      const uint32_t cnt = (uint32_t)c.reqs.size();
      int rc = XPG_OK;
      if (db->mx) {
        rc = batch_chunk(ctx, db, c.reqs);  // the matrix layout: one launch for any count
      } else if (!batch_pays(ctx, db, cnt)) {
        for (const xpg_request *r : c.reqs)
          if (!rc) rc = xpg_reply_dev(ctx, db, r->query_dev, r->reply_dev);
      } else if (ctx->batch_fuse) {
        fused.push_back(std::move(c));
      } else {
        rc = batch_chunk(ctx, db, c.reqs);
      }
      if (rc) return rc;
    }
  }
  return batch_fused(ctx, fused);
}

int xpg_reply(xpg_ctx *ctx, const xpg_db *db, const uint64_t *query, uint64_t query_words, uint64_t *reply,
              uint64_t reply_cap_words) {
  if (!ctx || !db || !query || !reply) return set_err(ctx, XPG_EINVAL, "null argument");
  const uint64_t qw = xpg_db_query_words(db);
  const uint64_t rw = (uint64_t)db->Pl[db->d] * 2 * db->prm.m * db->prm.n;
  if (query_words != qw) return set_err(ctx, XPG_ESIZE, "query size does not match the database grid");
  if (reply_cap_words < rw) return set_err(ctx, XPG_ESIZE, "reply buffer too small");
  (void)hipSetDevice(ctx->device);
  // device copies of the query and the reply in one context-owned buffer, kept across calls
  int rc = ensure_io_buf(ctx, (qw + rw) * 8);
  if (rc) return rc;
  uint64_t *qd = ctx->io_buf, *rd = ctx->io_buf + qw;
  rc = copy_h2d_staged(ctx, qd, query, qw * 8);
  if (!rc) rc = xpg_reply_dev(ctx, db, qd, rd);
  if (!rc && hipMemcpyAsync(reply, rd, rw * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
    rc = set_err(ctx, XPG_EDEVICE, "D2H copy");
  if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc) rc = set_err(ctx, XPG_EDEVICE, "sync");
  return rc;
}

// ---- timing ----
int xpg_timing_enable(xpg_ctx *ctx, int on) {
  if (!ctx) return XPG_EINVAL;
  ctx->timing = on == 2 ? 2 : on != 0 ? 1 : 0;
  return XPG_OK;
}

int xpg_timing_read(xpg_ctx *ctx, double *mac_ms, uint64_t *launches, int reset) {
  if (!ctx) return XPG_EINVAL;
  timing_collect(ctx);
  if (mac_ms) *mac_ms = ctx->mac_ms;
  if (launches) *launches = ctx->mac_launches;
  if (reset) {
    ctx->mac_ms = 0;
    ctx->mac_launches = 0;
  }
  return XPG_OK;
}

int xpg_graph_stats(xpg_ctx *ctx, uint64_t *captures, uint64_t *replays) {
  if (!ctx) return XPG_EINVAL;
  if (captures) *captures = ctx->graph_captures;
  if (replays) *replays = ctx->graph_launches;
  return XPG_OK;
}

int xpg_timing_read_batched(xpg_ctx *ctx, double *ms, uint64_t *launches, uint64_t *steps, int reset) {
  if (!ctx) return XPG_EINVAL;
  timing_collect(ctx);
  if (ms) *ms = ctx->batch_ms;
  if (launches) *launches = ctx->batch_launches;
  if (steps) *steps = ctx->batch_steps;
  if (reset) {
    ctx->batch_ms = 0;
    ctx->batch_launches = 0;
    ctx->batch_steps = 0;
  }
  return XPG_OK;
}

int xpg_timing_read_main(xpg_ctx *ctx, double *ms, uint64_t *launches, int reset) {
  if (!ctx) return XPG_EINVAL;
  timing_collect(ctx);
  if (ms) *ms = ctx->main_ms;
  if (launches) *launches = ctx->main_launches;
  if (reset) {
    ctx->main_ms = 0;
    ctx->main_launches = 0;
  }
  return XPG_OK;
}

// ---- client side ----
int xpg_keygen(const xpg_params *prm, uint64_t seed, uint64_t *sk) {
  if (!prm || !sk) return XPG_EINVAL;
  for (uint64_t c = 0; c < prm->m; c++) {
    const uint64_t p = prm->moduli[c];
    for (uint64_t i = 0; i < prm->n; i++) {
      uint64_t r = rand_word(seed, c * prm->n + i) & ((1ULL << 61) - 1);
      sk[c * prm->n + i] = r >= p ? r - p : r;
    }
  }
  return XPG_OK;
}

int xpg_query_gen_dev(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *sk_dev, uint32_t d,
                      const uint32_t *dims, uint64_t index, uint64_t noise_seed, uint64_t *query_dev) {
  if (!ctx || !prm || !sk_dev || !dims || !query_dev || d == 0 || d > XPG_MAX_LEVELS) return XPG_EINVAL;
  uint64_t x = index, elem = 0;
  const uint64_t mn = (uint64_t)prm->m * prm->n;
  for (uint32_t l = 0; l < d; l++) {
    EncSrc es{};
    es.sk = sk_dev;
    es.noise_seed = noise_seed;
    es.elem0 = elem;
    es.hot = x % dims[l];  // PIRQueryGenerator_internal.cpp:61-70
    x /= dims[l];
    int rc = launch_encrypt(ctx, prm, es, dims[l], query_dev + elem * 2 * mn);
    if (rc) return rc;
    elem += dims[l];
  }
  return XPG_OK;
}

int xpg_random_bytes_dev(xpg_ctx *ctx, uint64_t seed, uint64_t offset, uint8_t *dst_dev, uint64_t nbytes) {
  if (!ctx || !dst_dev) return XPG_EINVAL;
  return launch_random_bytes(ctx, seed, offset, dst_dev, nbytes);
}

}  // extern "C"
