// internal.hpp -- host-side structures shared by the HIP translation units of
// libxpir_amd.so.  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/xpir_amd.h"
#include "arith.hpp"

// NTT tables for one (n, m) on one device: per modulus n words each, the reference's
// layouts (crypto/NFLlib.cpp:44-194, crypto/NFLlib.hpp:571-586).
struct DevTables {
  uint32_t n = 0, m = 0, logn = 0;
  uint64_t *phi = nullptr, *phiq = nullptr;  // phi^j and Shoup quotients
  uint64_t *w = nullptr, *wq = nullptr;      // DIF twiddles, layer after layer (n-1 used)
  uint64_t *iw = nullptr, *iwq = nullptr;    // the same for the inverse root (decryption)
  uint64_t *ipt = nullptr, *iptq = nullptr;  // phi^-j * n^-1 and Shoup quotients
};

// Source of the coefficients an import kernel transforms.
enum SrcKind : int { SRC_BYTES = 0, SRC_ENCRYPT = 1 };

// Byte source: unit u covers bytes [u*stride, u*stride + len) of base, clipped at limit;
// bytes past the clip read as zero.
struct ByteSrc {
  const uint8_t *base;
  uint64_t stride, len, limit;
};

// Query-encryption source: element `elem0 + u` encrypts bit (u == hot) (NFLLWE.cpp:276-372).
struct EncSrc {
  const uint64_t *sk;  // m*n key words
  uint64_t noise_seed;
  uint64_t elem0;
  uint64_t hot;        // unit index whose message bit is 1 (or ~0)
  uint64_t emask, ebound;
  uint64_t amp[XPG_MAX_M], ampq[XPG_MAX_M];  // 2^A mod p and quotient
  uint64_t skq_unused;
};

// Level 0 in the matrix layout (mx.hip): S = m*n slots, U = groups*P units in UT tiles of
// 4, n_l elements in EB blocks of 32, of which the first E0 = n_l - 40 (in EBP blocks) run
// on the matrix cores and the last L = 40 with the reference's step.
struct MxGeom {
  uint32_t S = 0, U = 0, UT = 0, EB = 0, EBP = 0, L = 0, P = 0;
  uint64_t n_l = 0, groups = 0, E0 = 0;
};

struct xpg_db {
  // the importing context while it lives (xpg_ctx_destroy detaches its databases: a
  // database outlives its context and serves any context of the same GPU)
  xpg_ctx *ctx = nullptr;
  int device = 0;
  xpg_params prm{};
  uint32_t d = 1;
  uint32_t dims[XPG_MAX_LEVELS] = {0};
  uint64_t units = 0;      // prod(dims)
  uint64_t real_units = 0; // ceil(n_elems / alpha)
  uint32_t P0 = 0;
  uint32_t Pl[XPG_MAX_LEVELS + 1] = {0};  // polys per unit at each level; Pl[d] = reply polys
  uint64_t n_elems = 0, elem_bytes = 0;
  uint32_t alpha = 1;
  int alpha_zero_pad = 0;   // the partial-aggregate convention the import used
  uint64_t *words = nullptr;  // units x P0 x m x n (null in the matrix layout)
  // matrix layout (context option db_layout = 1 at import): level 0's digit fragments and
  // each group's last L elements row-major (the chains' tails); `words` is then freed
  uint8_t *mx = nullptr;
  uint64_t *mx_tail = nullptr;
  MxGeom mxg;
  uint64_t serial = 0;        // unique per database ever created (graph cache key)
};

// One captured reply (every launch of run_levels for one database, query and reply
// buffer), replayed with hipGraphLaunch.  Valid while the context's buffer epoch is
// unchanged: growing a context buffer or changing an option bumps the epoch.
struct GraphEntry {
  uint64_t db_serial = 0;
  const uint64_t *q = nullptr;
  uint64_t *r = nullptr;
  uint64_t epoch = 0;             // context epoch at capture
  uint64_t tune_gen = 0;          // tuning table generation at capture
  hipGraphExec_t exec = nullptr;  // null: seen once (ran eagerly); captured on the next call
  uint64_t last_use = 0;
};

struct xpg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::set<xpg_db *> dbs;  // live databases imported here
  std::string err;
  std::map<uint64_t, DevTables> tables;  // key (n << 8) | m
  // scratch reused across calls (grown on demand): block summaries of the streaming MAC
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  uint64_t *batch_buf = nullptr;  // level-0 replies of a batch (d > 1)
  size_t batch_bytes = 0;
  void *fused_buf = nullptr;      // tile counters and pointers of a fused level-0 launch
  size_t fused_bytes = 0;
  uint64_t *level_buf[2] = {nullptr, nullptr};
  size_t level_bytes[2] = {0, 0};
  // a call's workspace: while set, a reply's level buffers and block summaries and a
  // batch's intermediate levels are carved from it instead of the context's grown
  // buffers.  Set by xpg_reply_dev_noalloc (the device's static scratch array; then
  // nothing at all is allocated: noalloc) and by the serving loop (a worker's share of
  // the admitted pool).
  uint8_t *ws = nullptr;
  size_t ws_cap = 0, ws_used = 0;
  bool noalloc = false;
  uint64_t *io_buf = nullptr;     // device copies of a host query and its reply (xpg_reply)
  size_t io_bytes = 0;
  // two pinned host chunks through which large host buffers reach the device (host threads
  // fill one while the DMA engine drains the other), each with the event of its last copy
  void *pin[2] = {nullptr, nullptr};
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  // two small pinned buffers through which launch tables reach the device without the
  // stream drain a copy from pageable memory implies, taken in turn, each with the event of
  // its last copy
  void *tab_pin[2] = {nullptr, nullptr};
  size_t tab_bytes[2] = {0, 0};
  hipEvent_t tab_ev[2] = {nullptr, nullptr};
  int tab_next = 0;
  uint64_t *dec_buf = nullptr;    // residues and lifted coefficients of xpg_decrypt_dev
  size_t dec_bytes = 0;
  // MAC kernel choice: 0 = by shape, 1 = one chain per lane, 3 = streaming block summaries
  int mac_mode = 0;
  int stream_e = 0, stream_u = 0;  // streaming MAC: elements per block, per load chunk (0 = default)
  int stream_pf = 0;               // streaming MAC: double-buffered chunk loads
  int cu_count = 256;              // compute units of the device
  uint64_t group_budget = 0;       // bytes for a batch's grouped levels (0 = default)
  // batches: level 0 of every database chunk in fused persistent launches (1) or one
  // launch per chunk (0)
  int batch_fuse = 1;
  // a partial last alpha aggregate: 0 = the reference's bytes (readAggregatedStream's
  // bzero from the last real element's slot), 1 = the missing elements read as zeros
  int alpha_zero_pad = 0;
  // databases imported from now on: 0 = row-major NTT words, 1 = the matrix layout
  int db_layout = 0;
  // MAC timing
  int timing = 0;  // 0 off, 1 every MAC level and main kernel, 2 main kernels alone
  std::vector<hipEvent_t> ev_pool;
  // (start, end, 0 = whole MAC | 1 = main kernel | 2 = main kernel, register-blocked)
  std::vector<std::tuple<hipEvent_t, hipEvent_t, int>> pending;
  double mac_ms = 0.0;
  uint64_t mac_launches = 0;
  double main_ms = 0.0;        // the dominant kernel of each MAC (without its scan)
  uint64_t main_launches = 0;
  double batch_ms = 0.0;       // the register-blocked kernel's launches alone (tag 2)
  uint64_t batch_launches = 0, batch_steps = 0;  // ... and the exact chain steps they ran
  // captured replies (option "graphs")
  bool graphs = false;
  uint64_t epoch = 1;          // bumped whenever a context buffer moves or an option changes
  uint64_t use_clock = 0;
  std::vector<GraphEntry> graph_cache;
  uint64_t graph_launches = 0, graph_captures = 0;
};

// ---- host helpers (context.hip) ----
int set_err(xpg_ctx *ctx, int code, const std::string &msg);
const DevTables *get_tables(xpg_ctx *ctx, uint32_t n, uint32_t m);
int ensure_scratch(xpg_ctx *ctx, size_t bytes);
// scratch for one launch: the context's grown buffer, or during xpg_reply_dev_noalloc the
// free part of the static scratch array (nothing allocated; XPG_ESIZE if it is too small)
int get_scratch(xpg_ctx *ctx, size_t bytes, void **p);
int ensure_level_buf(xpg_ctx *ctx, int which, size_t bytes);
int ensure_batch_buf(xpg_ctx *ctx, size_t bytes);
int ensure_io_buf(xpg_ctx *ctx, size_t bytes);
int copy_h2d_staged(xpg_ctx *ctx, void *dst, const void *src, size_t bytes);
// a small host table to device memory through the context's pinned table buffers,
// ordered on its stream; src may be reused when the call returns
int copy_h2d_table(xpg_ctx *ctx, void *dst, const void *src, size_t bytes);
int ensure_dec_buf(xpg_ctx *ctx, size_t bytes);
void graph_cache_clear(xpg_ctx *ctx, uint64_t db_serial);  // 0 = every entry
ModSet modset_of(const xpg_params *prm);
void timing_begin(xpg_ctx *ctx, hipEvent_t *a, hipEvent_t *b, int tag = 0);
void timing_end(xpg_ctx *ctx, hipEvent_t a, hipEvent_t b, int tag = 0);

// Workspace bytes xpg_reply_batch_dev needs for k requests on db when given a workspace
// (the serving loop sizes its batches by it).
uint64_t batch_workspace_bytes(const xpg_ctx *ctx, const xpg_db *db, uint32_t k);
// xpg_reply_batch_dev with every intermediate buffer carved from [base, base + bytes)
int reply_batch_in_workspace(xpg_ctx *ctx, const xpg_request *reqs, uint32_t n_reqs, uint8_t *base, size_t bytes);

// ---- tuning (tune.hip) ----
// Kernel configuration of one MAC launch.  kernel: 1 = one chain per lane, 3 = streaming
// block summaries (a = elements per load chunk, b = double-buffered, c =
// elements per block, 0 = sized for one pass), 4 = register-blocked (a = groups per lane,
// b = queries per lane, c = 2: query words loaded per wave, 3: shared through LDS by the
// waves of a workgroup, 0: by shape).
struct LevelCfg {
  int kernel = 0;
  int a = 0, b = 0, c = 0;
};
struct TuneKey {
  uint32_t kind, n, m, P;
  uint64_t n_l, groups;
  uint32_t qn;
  bool operator<(const TuneKey &o) const {
    return std::tie(kind, n, m, P, n_l, groups, qn) < std::tie(o.kind, o.n, o.m, o.P, o.n_l, o.groups, o.qn);
  }
};
// process-wide table per device (every context on the device, the server's included)
bool tune_lookup(int device, const TuneKey &k, LevelCfg *cfg);
uint64_t tune_generation();
uint32_t tune_qn_bucket(uint32_t qn);  // batch sizes are tuned at powers of two

// ---- launchers (import.hip) ----
// Transform `units` units of a byte source into NTT words (units x P x m x n); the
// units in [units, units_total) are zero polys.
int launch_import_bytes(xpg_ctx *ctx, const xpg_params *prm, const ByteSrc &src, uint64_t units,
                        uint64_t units_total, uint32_t P, uint64_t *out);
// Encrypt `count` query elements (each a||b, canonical) into out.
int launch_encrypt(xpg_ctx *ctx, const xpg_params *prm, const EncSrc &src, uint64_t count,
                   uint64_t *out);
int launch_random_bytes(xpg_ctx *ctx, uint64_t seed, uint64_t offset, uint8_t *dst, uint64_t nbytes);
int launch_fill_zero(xpg_ctx *ctx, uint64_t *dst, uint64_t words);
// *out_dev = sum_i rand_word(words[i], i) mod 2^64 (the NTT file's checksum), stream ordered.
int launch_word_checksum(xpg_ctx *ctx, const uint64_t *words, uint64_t n, uint64_t *out_dev);

// ---- launchers (mac.hip) ----
// One reply level: out (groups x P x 2 x m x n) from db (groups*n_l x P x m x n) and
// q (n_l x 2 x m x n).
int launch_mac_level(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *db, const uint64_t *q,
                     uint64_t n_l, uint64_t groups, uint32_t P, uint64_t *out);
// QN (1..64) queries over one database level; q / out are host arrays of QN device
// pointers (they travel in the kernel arguments)
int launch_mac_batch(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *db, const uint64_t *const *q,
                     uint64_t *const *out, uint32_t qn, uint64_t n_l, uint64_t groups, uint32_t P);
// The same with an explicit configuration (the tuner's candidates): launch_mac_level_cfg
// takes kernels 1..4, launch_mac_batch_cfg kernel 4 (a x b register blocks, 0 = by shape).
int launch_mac_level_cfg(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *db, const uint64_t *q,
                         uint64_t n_l, uint64_t groups, uint32_t P, uint64_t *out, const LevelCfg &cfg);
int launch_mac_batch_cfg(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *db, const uint64_t *const *q,
                         uint64_t *const *out, uint32_t qn, uint64_t n_l, uint64_t groups, uint32_t P, int gr,
                         int qr, int family);
// Level 0 of several database chunks in one persistent launch (mac_fused_kernel): part i
// is qn (1..64) queries over the level-0 database db (n_l elements per chain, `groups`
// groups of P polys); q / out are host arrays of qn device pointers.  Every part's moduli
// are the first part's first m.  dev_table is device memory of fused_table_bytes(total
// queries) bytes that the launch uploads its tile counters and pointers to (stream
// ordered: it may be reused by the next launch on the stream).  A chunk takes
// fused_segments(qn) of the launch's kFusedMaxSegments segments.
struct FusedPart {
  const xpg_params *prm;
  const uint64_t *db;
  const uint64_t *const *q;
  uint64_t *const *out;
  uint32_t qn;
  uint64_t n_l, groups;
  uint32_t P;
};
constexpr uint32_t kFusedMaxSegments = 24;
size_t fused_table_bytes(uint32_t queries);
uint32_t fused_segments(uint32_t qn);
int launch_mac_fused(xpg_ctx *ctx, const FusedPart *parts, uint32_t nparts, void *dev_table);
// One level of up to 64 requests at once: their level databases side by side (request i's
// groups are groups [i*gpr, (i+1)*gpr) of db), request i reading query q[i]; outputs to
// out_per_req[i] (each request's own buffer) when given, else contiguous in out.
int launch_mac_level_multi(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *db, const uint64_t *const *q,
                           uint32_t nreq, uint64_t n_l, uint64_t gpr, uint32_t P, uint64_t *const *out_per_req,
                           uint64_t *out);
// (mac.hip) The last L steps of level-0 chains continued from residues the matrix-core
// level 0 computed (mx.hip): tail_db holds each group's last L elements (groups*L rows x
// P x m x n), q[i] query i's first tail element; the words go to out.
// The chains' starting residues come from the matrix kernel's staging buffer when
// stage.stage is set (items of cb column blocks x rb row blocks, each wave holding mt row
// tiles; qfirst: the index of q[0] in the staged launch), else from out.
struct TailStage {
  const uint64_t *stage = nullptr;
  uint32_t cb = 0, rb = 0, mt = 0;
};
int launch_mac_tail(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *tail_db, const uint64_t *const *q,
                    uint64_t *const *out, uint32_t qn, uint64_t L, uint64_t groups, uint32_t P, const TailStage &stage,
                    uint32_t qfirst);

// ---- matrix-core level 0 (mx.hip) ----
// Whether level 0 of (n_l, groups, P) can take the matrix layout (n a multiple of 64, m*n
// of 256, 72 <= n_l < 2^27), and its geometry.
bool mx_geometry(const xpg_params *prm, uint64_t n_l, uint64_t groups, uint32_t P, MxGeom *g);
uint64_t mx_frag_bytes(const MxGeom &g);
uint64_t mx_tail_words(const MxGeom &g);
// The fragments of row-major level-0 words and the tails' rows; *bad_dev (zero before the
// launch) becomes nonzero when a word is at least 2^63 - 2^56 (no 8-digit form).
int launch_mx_pack_db(xpg_ctx *ctx, const xpg_params *prm, const MxGeom &g, const uint64_t *words, uint8_t *frag,
                      uint64_t *tail, uint32_t *bad_dev);
// device workspace of one launch of qn (1..64) queries
uint64_t mx_work_bytes(const MxGeom &g, uint32_t qn);
// Level 0 of qn queries (q[i]: level-0 query words; out[i]: groups x P x 2 x m x n) on a
// database in the matrix layout, the same words as launch_mac_level / launch_mac_batch.
int launch_mx_level(xpg_ctx *ctx, const xpg_params *prm, const MxGeom &g, const uint8_t *frag, const uint64_t *tail,
                    const uint64_t *const *q, uint64_t *const *out, uint32_t qn, void *work);

// what launch_mac_level runs for this shape: the tuning table's choice unless an option
// pins a kernel, else default_level_cfg
LevelCfg resolve_level_cfg(const xpg_ctx *ctx, const xpg_params *prm, uint64_t n_l, uint64_t groups, uint32_t P);
// device scratch a level launch with cfg takes (the streaming kernel's block summaries)
uint64_t level_scratch_bytes(const xpg_ctx *ctx, const xpg_params *prm, uint64_t n_l, uint64_t groups, uint32_t P,
                             const LevelCfg &cfg);
// what launch_mac_level runs for this shape with the context's options and no tuning
LevelCfg default_level_cfg(const xpg_ctx *ctx, const xpg_params *prm, uint64_t n_l, uint64_t groups, uint32_t P);
bool batch_shape_built(int GR, int QR);   // register-block shapes compiled in (per-wave query loads)
bool batch3_shape_built(int GR, int QR);  // ... and with the query words shared through LDS
bool batch_family_built(int family, int GR, int QR);  // family 2 or 3 (anything else: none)
// explicit options (a kernel or a tile shape asked for) win over the tuning table
bool options_pinned(const xpg_ctx *ctx);
// register blocks the batched MAC picks by shape for the ones given as 0
void batch_default_shape(uint32_t qn, uint64_t groups, int *gr, int *qr);
// stream E that launch_stream_level picks when none is given
uint64_t stream_auto_e(const xpg_ctx *ctx, const xpg_params *prm, uint64_t n_l, uint64_t groups, uint32_t P, int u);

