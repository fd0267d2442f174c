// mac.hip -- the reply MAC on CDNA4: Shoup products of NTT-domain database words with
// query words, accumulated per chain with the reference's lazy step, in element order.
//
// Reference path: PIRReplyGeneratorNFL_internal::generateReply(queries, data, lvl, result)
// (pir/replyGenerator/PIRReplyGeneratorNFL_internal.cpp:511-595) -> NFLLWE::mulandadd
// (Shoup form, crypto/NFLLWE.cpp:176-205) -> NFLlib::mulandaddShoup (crypto/NFLlib.hpp:303-321).
//
// A chain is (group g, poly k, modulus c, slot s, component a|b); its accumulator takes
// one step per element e = 0..n_l-1 in order:
//     acc = acc + (x*y - hi64(x*y')*p);  if (acc >= p) acc -= p
// with x the lazy DB word, y the canonical query word and y' = floor(y*2^64/p).
// The final words are not canonical (about 0.4% lie in [p,2p)), so the chain order is
// part of the result.  y' is derived in registers from the special form of p (never
// stored: for a single query it would double the query bytes).
#include <algorithm>

#include "internal.hpp"

namespace {

struct MacArgs {
  ModSet ms;
  const uint64_t *db;
  const uint64_t *q;
  uint64_t *out;
  uint64_t n_l, groups;
  uint32_t P, n, m, logn;
};

// Both chains (a and b) of one lane over n_l elements, in element order: x from the
// database column, ya / yb from the query's a and b halves (strides in words).
__device__ __forceinline__ void direct_chains(const uint64_t *x, const uint64_t *ya, const uint64_t *yb,
                                              uint64_t n_l, uint64_t xstride, uint64_t qstride, uint64_t p,
                                              uint32_t delta, uint64_t &acc_a, uint64_t &acc_b) {
  acc_a = acc_b = 0;
  const StepConsts kc = step_consts(p, delta);
  uint64_t e = 0;
  for (; e + 4 <= n_l; e += 4) {
    uint64_t xv[4], av[4], bv[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      xv[i] = __builtin_nontemporal_load(x + (e + i) * xstride);
      av[i] = ya[(e + i) * qstride];
      bv[i] = yb[(e + i) * qstride];
    }
    __builtin_amdgcn_sched_barrier(0);  // the chunk's loads stay ahead of its arithmetic
#pragma unroll
    for (int i = 0; i < 4; i++) {
      mac_step_asm(acc_a, xv[i], av[i], shoup_quot(av[i], p, delta), kc);
      mac_step_asm(acc_b, xv[i], bv[i], shoup_quot(bv[i], p, delta), kc);
    }
  }
  for (; e < n_l; e++) {
    const uint64_t xv = x[e * xstride], av = ya[e * qstride], bv = yb[e * qstride];
    mac_step_asm(acc_a, xv, av, shoup_quot(av, p, delta), kc);
    mac_step_asm(acc_b, xv, bv, shoup_quot(bv, p, delta), kc);
  }
}

// One lane = both chains (a and b) of one (g, k, c, s): x is loaded once per element.
// Lanes with consecutive s read consecutive words, so every load is a coalesced 512 B
// row per wavefront.  The query quotients are derived in registers.
__global__ void __launch_bounds__(256) mac_direct_kernel(MacArgs a) {
  const uint64_t mn = (uint64_t)a.m * a.n;
  const uint64_t lane = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t total = a.groups * a.P * mn;
  if (lane >= total) return;
  const uint64_t cs = lane % mn;          // c*n + s
  const uint32_t c = (uint32_t)(cs >> a.logn);
  const uint64_t gk = lane / mn;
  const uint32_t k = (uint32_t)(gk % a.P);
  const uint64_t g = gk / a.P;
  const uint64_t xstride = (uint64_t)a.P * mn;
  uint64_t acc_a, acc_b;
  direct_chains(a.db + (g * a.n_l * a.P + k) * mn + cs, a.q + cs, a.q + mn + cs, a.n_l, xstride, 2 * mn,
                a.ms.p[c], a.ms.delta[c], acc_a, acc_b);
  uint64_t *o = a.out + gk * 2 * mn + cs;
  o[0] = acc_a;
  o[mn] = acc_b;
}

// Several requests' levels in one launch: group G of the launch belongs to request
// G / gpr and reads that request's query (q[req]); its output goes to out[req] (the
// request's own buffer, group G mod gpr) when per_req_out, else to base.out + G.
constexpr int kMaxReqs = 64;
struct MultiArgs {
  MacArgs base;
  const uint64_t *q[kMaxReqs];
  uint64_t *out[kMaxReqs];
  uint64_t gpr;          // groups per request
  uint32_t per_req_out;
};

__global__ void __launch_bounds__(256) mac_direct_multi_kernel(MultiArgs m) {
  const MacArgs &a = m.base;
  const uint64_t mn = (uint64_t)a.m * a.n;
  const uint64_t lane = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t total = a.groups * a.P * mn;
  if (lane >= total) return;
  const uint64_t cs = lane % mn;
  const uint32_t c = (uint32_t)(cs >> a.logn);
  const uint64_t gk = lane / mn;
  const uint32_t k = (uint32_t)(gk % a.P);
  const uint64_t g = gk / a.P;
  const uint64_t req = g / m.gpr;
  const uint64_t *q = m.q[req];
  uint64_t acc_a, acc_b;
  direct_chains(a.db + (g * a.n_l * a.P + k) * mn + cs, q + cs, q + mn + cs, a.n_l, (uint64_t)a.P * mn, 2 * mn,
                a.ms.p[c], a.ms.delta[c], acc_a, acc_b);
  uint64_t *o = m.per_req_out ? m.out[req] + (gk - req * m.gpr * a.P) * 2 * mn + cs : a.out + gk * 2 * mn + cs;
  o[0] = acc_a;
  o[mn] = acc_b;
}

// The same launch with one lane per (group, slot) carrying all PR polys of the group: the
// element's query words are loaded and their Shoup quotients derived once for the PR polys
// (one chain per lane reloads them per poly); the next element's words load while this
// one's steps run.  Each chain still takes its elements in order through the same step.
template <int PR>
__global__ void __launch_bounds__(256) mac_multi_rows_kernel(MultiArgs m) {
  const MacArgs &a = m.base;
  const uint64_t mn = (uint64_t)a.m * a.n;
  const uint64_t lane = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (lane >= a.groups * mn) return;
  const uint64_t cs = lane % mn;
  const uint32_t c = (uint32_t)(cs >> a.logn);
  const uint64_t g = lane / mn;
  const uint64_t req = g / m.gpr;
  const uint64_t *q = m.q[req] + cs;
  const uint64_t p = a.ms.p[c];
  const uint32_t delta = a.ms.delta[c];
  const StepConsts kc = step_consts(p, delta);
  const uint64_t *x = a.db + g * a.n_l * PR * mn + cs;  // element e, poly k at (e * PR + k) * mn
  uint64_t acc_a[PR], acc_b[PR];
#pragma unroll
  for (int k = 0; k < PR; k++) acc_a[k] = acc_b[k] = 0;
  uint64_t xv[2][PR], ya[2], yb[2];
  auto load = [&](int b, uint64_t e) {
#pragma unroll
    for (int k = 0; k < PR; k++) xv[b][k] = __builtin_nontemporal_load(x + (e * PR + k) * mn);
    ya[b] = q[e * 2 * mn];
    yb[b] = q[e * 2 * mn + mn];
  };
  auto steps = [&](int b) {
    const uint64_t yaq = shoup_quot(ya[b], p, delta), ybq = shoup_quot(yb[b], p, delta);
#pragma unroll
    for (int k = 0; k < PR; k++) {
      mac_step_asm(acc_a[k], xv[b][k], ya[b], yaq, kc);
      mac_step_asm(acc_b[k], xv[b][k], yb[b], ybq, kc);
    }
  };
  const uint64_t nl = a.n_l;
  load(0, 0);
  uint64_t e = 0;
  for (; e + 2 <= nl; e += 2) {
    load(1, e + 1);
    steps(0);
    if (e + 2 < nl) load(0, e + 2);
    steps(1);
  }
  if (e < nl) steps(0);
  const uint64_t first = (g - req * m.gpr) * PR;
#pragma unroll
  for (int k = 0; k < PR; k++) {
    uint64_t *o = m.per_req_out ? m.out[req] + (first + k) * 2 * mn : a.out + (g * PR + k) * 2 * mn;
    o[cs] = acc_a[k];
    o[cs + mn] = acc_b[k];
  }
}

// ------------------------------------------------------------------------------------
// Levels with several groups, one or many queries.
constexpr int kMaxBatch = 64;  // queries per launch; pointers travel in the kernel arguments

// Register-blocked level: a lane carries GR groups x QR queries (x a|b) of one (k, c, s).
// Per element it loads GR DB words and QR query word pairs, derives the pairs' Shoup
// quotients once, then does 2*GR*QR Shoup steps, reusing each DB word QR times and each
// query word (and its quotient) GR times from registers.  Levels with many groups are
// bound by how often query words are re-read from L2/MALL across groups, which this
// blocking divides by GR.  The query blocks of one column group run on one XCD, so
// its DB rows come from HBM once and from L2 for the other query blocks.
struct Batch2Args {
  ModSet ms;
  const uint64_t *db;
  const uint64_t *q[kMaxBatch];
  uint64_t *out[kMaxBatch];
  uint64_t n_l, groups, colgroups;  // colgroups = row blocks * m * sblocks
  uint32_t P, n, m, sblocks, qn, qblocks;
  // tail chains: their starting residues in the matrix kernel's staging (null: in out)
  const uint64_t *stage;
  uint32_t st_cb, st_rb, st_mt, qfirst;  // qfirst: the launch's first query in the chunk
};

// Word at byte offset `lane_off` from a wave-uniform base: written this way the load
// takes the uniform base in SGPRs and the lane's 32-bit offset in a VGPR, so no per-load
// VALU address arithmetic is needed.
__device__ __forceinline__ const uint64_t *lane_word(const uint64_t *base, uint32_t lane_off) {
  return reinterpret_cast<const uint64_t *>(reinterpret_cast<const char *>(base) + lane_off);
}

// Block b -> (colgroup cg, query block qb).  A colgroup is (row block, column), the column
// (c, sb) fastest; the 8 XCDs take blocks round-robin (b & 7) and the query blocks of one
// colgroup run back to back on one XCD, so its DB rows come from HBM once and from that
// XCD's L2 for the other query blocks.
__device__ __forceinline__ bool batch_tile(const Batch2Args &a, uint64_t b, uint64_t &cg, uint32_t &qb) {
  const uint64_t xcd = b & 7, j = b >> 3;
  qb = (uint32_t)(j % a.qblocks);
  cg = (j / a.qblocks) * 8 + xcd;
  return cg < a.colgroups;
}

template <int GR, int QR>
__global__ void __launch_bounds__(64) mac_batch2_kernel(Batch2Args a) {
  uint64_t cg;
  uint32_t qb;
  if (!batch_tile(a, blockIdx.x, cg, qb)) return;
  const uint32_t lane = threadIdx.x;
  const uint32_t sb = (uint32_t)(cg % a.sblocks);
  uint64_t r = cg / a.sblocks;
  const uint32_t c = (uint32_t)(r % a.m);
  r /= a.m;
  const uint32_t k = (uint32_t)(r % a.P);
  const uint64_t g0 = (r / a.P) * GR;
  const uint64_t mn = (uint64_t)a.m * a.n;
  const uint64_t cs = (uint64_t)c * a.n + sb * 64 + lane;
  const uint64_t p = a.ms.p[c];
  const uint64_t xstride = (uint64_t)a.P * mn, qstride = 2 * mn, gstride = a.n_l * xstride;
  // wave-uniform row and query bases; the lane adds its byte offset cs8 at each load
  const uint64_t *x = a.db + (g0 * a.n_l * a.P + k) * mn;
  const uint32_t cs8 = (uint32_t)cs * 8;
  const uint64_t *qa[QR];
  const uint32_t q0 = qb * QR;
#pragma unroll
  for (int i = 0; i < QR; i++) qa[i] = a.q[min(q0 + i, a.qn - 1)];
  const uint32_t delta = a.ms.delta[c];
  const StepConsts kc = step_consts(p, delta);
  int gv = (int)min<uint64_t>(GR, a.groups - g0);  // groups of this block that exist
  uint64_t acc_a[GR][QR], acc_b[GR][QR];
#pragma unroll
  for (int g = 0; g < GR; g++)
#pragma unroll
    for (int i = 0; i < QR; i++) acc_a[g][i] = acc_b[g][i] = 0;
  // UE elements per iteration: every load of the chunk is issued before its arithmetic,
  // which then runs in element order (the chains' order)
  constexpr int UE = GR * QR >= 16 ? 1 : 4;
  const uint32_t nl = (uint32_t)a.n_l;  // element indices in 32 bits: scalar arithmetic
  uint32_t e = 0;
  if (gv == GR) {
    // every group of the block exists: whole chunks without per-load guards
    for (; e + UE <= nl; e += UE) {
      uint64_t xv[UE][GR], ya[UE][QR], yb[UE][QR];
#pragma unroll
      for (int u = 0; u < UE; u++) {
        const uint64_t eo = (uint64_t)(e + u) * xstride, o = (uint64_t)(e + u) * qstride;
#pragma unroll
        for (int g = 0; g < GR; g++) xv[u][g] = __builtin_nontemporal_load(lane_word(x + g * gstride + eo, cs8));
#pragma unroll
        for (int i = 0; i < QR; i++) {
          ya[u][i] = *lane_word(qa[i] + o, cs8);
          yb[u][i] = *lane_word(qa[i] + o + mn, cs8);
        }
      }
      // keep the whole chunk's loads ahead of its arithmetic (the scheduler would
      // otherwise sink each load to its first use and wait for it there)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < UE; u++) {
#pragma unroll
        for (int i = 0; i < QR; i++) {
          const uint64_t yaq = shoup_quot(ya[u][i], p, delta), ybq = shoup_quot(yb[u][i], p, delta);
#pragma unroll
          for (int g = 0; g < GR; g++) {
            mac_step_asm(acc_a[g][i], xv[u][g], ya[u][i], yaq, kc);
            mac_step_asm(acc_b[g][i], xv[u][g], yb[u][i], ybq, kc);
          }
        }
      }
    }
  }
  // the tail chunk, and blocks whose last groups lie past the level
  for (; e < nl; e += UE) {
    const int ve = (int)min<uint32_t>(UE, nl - e);
    uint64_t xv[UE][GR], ya[UE][QR], yb[UE][QR];
#pragma unroll
    for (int u = 0; u < UE; u++) {
      const bool in = u < ve;
      const uint64_t eo = (uint64_t)(e + u) * xstride, o = (uint64_t)(e + u) * qstride;
#pragma unroll
      for (int g = 0; g < GR; g++)
        xv[u][g] = in && g < gv ? __builtin_nontemporal_load(lane_word(x + g * gstride + eo, cs8)) : 0;
#pragma unroll
      for (int i = 0; i < QR; i++) {
        ya[u][i] = in ? *lane_word(qa[i] + o, cs8) : 0;
        yb[u][i] = in ? *lane_word(qa[i] + o + mn, cs8) : 0;
      }
    }
#pragma unroll
    for (int u = 0; u < UE; u++) {
      if (u < ve) {
#pragma unroll
        for (int i = 0; i < QR; i++) {
          const uint64_t yaq = shoup_quot(ya[u][i], p, delta), ybq = shoup_quot(yb[u][i], p, delta);
#pragma unroll
          for (int g = 0; g < GR; g++) {
            mac_step_asm(acc_a[g][i], xv[u][g], ya[u][i], yaq, kc);
            mac_step_asm(acc_b[g][i], xv[u][g], yb[u][i], ybq, kc);
          }
        }
      }
    }
  }
#pragma unroll
  for (int g = 0; g < GR; g++) {
    if (g >= gv) continue;
    const uint64_t gk = (g0 + g) * a.P + k;
#pragma unroll
    for (int i = 0; i < QR; i++) {
      if (q0 + i < a.qn) {
        uint64_t *o = a.out[q0 + i] + gk * 2 * mn + cs;
        o[0] = acc_a[g][i];
        o[mn] = acc_b[g][i];
      }
    }
  }
}

// ------------------------------------------------------------------------------------
// Register blocks with the query words shared through LDS.  A workgroup of W = 4 waves
// takes one 64-slot column (modulus c, slots sb*64..+63), one block of QR queries and
// W*GR rows, a row being one (group g, poly k) of the level (row r = g*P + k);
// lane = slot.  Every row of a column reads the same query words, so the workgroup loads
// each query word once per chunk of kB3U elements, derives its Shoup quotient once, and
// stages (y, y') in LDS for its waves; each wave then runs the a and b chains of its GR
// rows x QR queries for the chunk, element by element, with the loads of the next chunk
// (query and DB words) in flight.  Two LDS buffers alternate between chunks; one barrier
// per chunk.  Same words as every other MAC kernel: each chain still takes its elements in
// order through the same step.
constexpr int kB3U = 2;  // elements per staged chunk

struct QPair {
  uint64_t y, yq;
};

// One level's database as a tile sees it (the per-database launch takes it from its
// arguments, the fused launch from its segment table).
struct TileLevel {
  const uint64_t *db;
  uint64_t n_l, groups;
  uint32_t P, n, m, sblocks, qn;
  // TAIL: the matrix kernel's staged residues (items of CB column x RB row blocks, MT row
  // tiles per wave) and the tile's first query; null: the residues are in the outputs
  const uint64_t *stage = nullptr;
  uint32_t st_cb = 0, st_rb = 0, st_mt = 0, q0 = 0;
};

// The tile (colgroup cg = (row block, column), query block qb) of one level: every wave
// of the workgroup runs it; qbase[i] / obase[i] are the query and output of the block's
// query i (the last query repeated past qn: its chains are computed and dropped).
// The word a tail chain leaves after steps of which `kept` did not subtract p.  Written
// acc = rho + j*p, a step maps j to max(j + w, 0) (w in {-1, 0, 1}); over the tail, a start
// at j0 ends at max(j0 + D, M) and the start at 0 at M, with D = M - kept, so every start
// j0 <= kept ends where the start at 0 does.  The chain's true start has j0 <= 8 (its word
// is below 2^64 < 9p), so kept >= 8 settles it; otherwise bit 63 (never set in a chain
// word, which stays below 2^62) asks for the whole chain.
__device__ __forceinline__ uint64_t tail_word(uint64_t acc, uint32_t kept) {
  return kept >= 8 ? acc : acc | (1ull << 63);
}

// TAIL: the chains continue from the words already in the outputs instead of zero, and a
// chain whose steps do not show that its start's excess multiple of p is forgotten gets
// bit 63 set in its word (for a whole recomputation): see the matrix-core level 0, whose
// chains' last steps this runs.
template <int GR, int QR, int W, bool TAIL, int NQMAX>
__device__ __forceinline__ void batch3_tile(const ModSet &ms, const TileLevel &a, const uint64_t *const (&qbase)[QR],
                                            uint64_t *const (&obase)[QR], uint32_t nq, uint64_t cg,
                                            QPair (&lds)[2][NQMAX][64]) {
  constexpr int NQ = kB3U * QR * 2;  // staged query rows (element, query, a|b) per chunk
  constexpr int NS = NQ / W;         // rows staged per wave
  static_assert(NQ % W == 0, "staged rows split evenly over the waves");
  static_assert(NQ <= NQMAX, "staging buffer holds the block's rows");
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t sb = (uint32_t)(cg % a.sblocks);
  uint64_t rr = cg / a.sblocks;
  const uint32_t c = (uint32_t)(rr % a.m);
  const uint64_t rb = rr / a.m;
  const uint64_t rows = a.groups * a.P;
  const uint64_t mn = (uint64_t)a.m * a.n;
  const uint32_t cs = c * a.n + sb * 64 + lane;  // this lane's word within a poly
  const uint64_t p = ms.p[c];
  const uint32_t delta = ms.delta[c];
  const StepConsts kc = step_consts(p, delta);
  const uint64_t xstride = (uint64_t)a.P * mn;  // next element of a row
  const uint64_t qstride = 2 * mn;              // next element of a query
  const uint64_t r0 = (rb * W + w) * GR;
  const int rv = (int)max<int64_t>(0, min<int64_t>(GR, (int64_t)rows - (int64_t)r0));  // this wave's rows
  // wave-uniform row bases (rows past the level read the last row: their chains are
  // computed and dropped, so no load needs a guard); the lane adds its 32-bit word offset
  const uint64_t *xr[GR];
#pragma unroll
  for (int g = 0; g < GR; g++) {
    const uint64_t r = min<uint64_t>(r0 + g, rows - 1);
    xr[g] = a.db + ((r / a.P) * a.n_l * a.P + (r % a.P)) * mn;
  }
  // staging: this wave loads the query rows qrow = w + W * s of every chunk
  const uint64_t *sq[NS];
  uint32_t su[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const uint32_t qrow = w + W * s;
    const uint32_t u = qrow / (QR * 2), rem = qrow % (QR * 2), i = rem >> 1, ab = rem & 1;
    su[s] = u;
    sq[s] = qbase[i] + ab * mn;
  }
  uint64_t acc_a[GR][QR], acc_b[GR][QR];
  uint32_t ns_a[GR][QR], ns_b[GR][QR];  // TAIL: steps that subtracted p
#pragma unroll
  for (int g = 0; g < GR; g++)
#pragma unroll
    for (int i = 0; i < QR; i++) {
      acc_a[g][i] = acc_b[g][i] = 0;
      ns_a[g][i] = ns_b[g][i] = 0;
      if (TAIL && g < rv && (uint32_t)i < nq) {
        if (a.stage) {
          // the residue of (slot cs, unit u, query q, half) in its staged item: item
          // (slot, unit block of 32, query block of 2 MT), within it wave (u % 32) / 8,
          // row tile (q / 2) % MT, column tile (u % 8) / 4, row group (q % 2) * 2 + half,
          // unit u % 4 (mx.hip: the matrix kernel's epilogue)
          const uint32_t u = (uint32_t)(r0 + g), q = a.q0 + i, mt = a.st_mt;
          const uint64_t item = ((uint64_t)cs * a.st_cb + u / 32) * a.st_rb + (q / 2) / mt;
          const uint64_t w = item * (mt * 128) + ((u % 32) / 8) * (mt * 32) + (((q / 2) % mt) * 2 + (u % 8) / 4) * 16 +
                             (q % 2) * 8 + (u % 4);
          acc_a[g][i] = a.stage[w];
          acc_b[g][i] = a.stage[w + 4];
        } else {
          const uint64_t *o = obase[i] + (r0 + g) * 2 * mn + cs;
          acc_a[g][i] = o[0];
          acc_b[g][i] = o[mn];
        }
      }
    }
  auto step2 = [&](int g, int i, uint64_t x, uint64_t ya, uint64_t yaq, uint64_t yb, uint64_t ybq) {
    if constexpr (TAIL) {
      mac_step_count_asm(acc_a[g][i], ns_a[g][i], x, ya, yaq, kc);
      mac_step_count_asm(acc_b[g][i], ns_b[g][i], x, yb, ybq, kc);
    } else {
      mac_step_asm(acc_a[g][i], x, ya, yaq, kc);
      mac_step_asm(acc_b[g][i], x, yb, ybq, kc);
    }
  };
  // element indices in 32 bits, so that the clamps and offsets below stay scalar
  const uint32_t nl = (uint32_t)a.n_l, nfull = nl / kB3U, last = nl - 1;
  const uint32_t cs8 = cs * 8;  // this lane's byte offset within a row

  // loads of chunk ch, element indices clamped to the level (a chunk past the last whole
  // one reads words it never uses)
  auto load_q = [&](uint32_t ch, uint64_t (&qv)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; s++)
      qv[s] = __builtin_nontemporal_load(lane_word(sq[s] + (uint64_t)min(ch * kB3U + su[s], last) * qstride, cs8));
  };
  auto load_x = [&](uint32_t ch, uint64_t (&xv)[kB3U][GR]) {
#pragma unroll
    for (int u = 0; u < kB3U; u++) {
      const uint64_t eo = (uint64_t)min(ch * kB3U + u, last) * xstride;
#pragma unroll
      for (int g = 0; g < GR; g++) xv[u][g] = __builtin_nontemporal_load(lane_word(xr[g] + eo, cs8));
    }
  };
  auto stage = [&](int buf, const uint64_t (&qv)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; s++) lds[buf][w + W * s][lane] = QPair{qv[s], shoup_quot(qv[s], p, delta)};
  };
  auto compute = [&](int buf, const uint64_t (&xv)[kB3U][GR]) {
#pragma unroll
    for (int u = 0; u < kB3U; u++)
#pragma unroll
      for (int i = 0; i < QR; i++) {
        const QPair ya = lds[buf][(u * QR + i) * 2][lane], yb = lds[buf][(u * QR + i) * 2 + 1][lane];
#pragma unroll
        for (int g = 0; g < GR; g++) step2(g, i, xv[u][g], ya.y, ya.yq, yb.y, yb.yq);
      }
  };

  uint64_t qv[NS], x0[kB3U][GR], x1[kB3U][GR];
  load_q(0, qv);
  load_x(0, x0);
  stage(0, qv);
  __syncthreads();
  for (uint32_t ch = 0; ch < nfull; ch += 2) {
    // whole chunk ch from buffer 0 and x0 while chunk ch + 1 loads, then chunk ch + 1 from
    // buffer 1 and x1 (when it is whole) while chunk ch + 2 loads
    load_q(ch + 1, qv);
    load_x(ch + 1, x1);
    compute(0, x0);
    stage(1, qv);
    __syncthreads();
    load_q(ch + 2, qv);
    load_x(ch + 2, x0);
    if (ch + 1 < nfull) compute(1, x1);
    stage(0, qv);
    __syncthreads();
  }
  // the elements after the last whole chunk: each wave loads its own query words
  for (uint32_t e = nfull * kB3U; e < nl; e++) {
    uint64_t xv[GR];
#pragma unroll
    for (int g = 0; g < GR; g++) xv[g] = __builtin_nontemporal_load(lane_word(xr[g] + (uint64_t)e * xstride, cs8));
#pragma unroll
    for (int i = 0; i < QR; i++) {
      const uint64_t *qi = qbase[i] + (uint64_t)e * qstride;
      const uint64_t ya = *lane_word(qi, cs8), yb = *lane_word(qi + mn, cs8);
      const uint64_t yaq = shoup_quot(ya, p, delta), ybq = shoup_quot(yb, p, delta);
#pragma unroll
      for (int g = 0; g < GR; g++) step2(g, i, xv[g], ya, yaq, yb, ybq);
    }
  }
#pragma unroll
  for (int g = 0; g < GR; g++) {
    if (g >= rv) continue;
#pragma unroll
    for (int i = 0; i < QR; i++) {
      if ((uint32_t)i < nq) {
        uint64_t *o = obase[i] + (r0 + g) * 2 * mn + cs;
        if constexpr (TAIL) {
          o[0] = tail_word(acc_a[g][i], nl - ns_a[g][i]);
          o[mn] = tail_word(acc_b[g][i], nl - ns_b[g][i]);
        } else {
          o[0] = acc_a[g][i];
          o[mn] = acc_b[g][i];
        }
      }
    }
  }
}

template <int GR, int QR, int W, bool TAIL = false>
__global__ void __launch_bounds__(64 * W) mac_batch3_kernel(Batch2Args a) {
  __shared__ QPair lds[2][kB3U * QR * 2][64];
  uint64_t cg;
  uint32_t qb;
  if (!batch_tile(a, blockIdx.x, cg, qb)) return;  // the whole workgroup
  const uint32_t q0 = qb * QR;
  const TileLevel lv{a.db,    a.n_l,    a.groups, a.P,     a.n,     a.m,           a.sblocks,
                     a.qn,    a.stage,  a.st_cb,  a.st_rb, a.st_mt, a.qfirst + q0};
  const uint64_t *qbase[QR];
  uint64_t *obase[QR];
#pragma unroll
  for (int i = 0; i < QR; i++) {
    qbase[i] = a.q[min(q0 + i, a.qn - 1)];
    obase[i] = a.out[min(q0 + i, a.qn - 1)];
  }
  batch3_tile<GR, QR, W, TAIL>(a.ms, lv, qbase, obase, a.qn - q0, cg, lds);
}

// ------------------------------------------------------------------------------------
// Level 0 of many databases in one persistent launch.  The launch's work is a table of
// segments, one per (database chunk, query block width): a segment is the register-
// blocked tiles (colgroup, query block) of one database's level 0 for its queries, in
// blocks of QR = 4 queries and one block of the 1..3 left over.  Tiles are dealt to the
// eight XCDs by colgroup (cg mod 8, as in the per-database launch, so the query blocks of
// one colgroup share that XCD's L2), and each XCD's list takes the segments in the
// order the host gives them (longest chains first).  The grid holds as many workgroups
// as are resident at once; each takes tiles from its own XCD's list through an atomic
// counter, and when that list is empty from the other XCDs' lists in turn, until every
// list is empty (the only exit; no workgroup waits for another).  The words are the
// same as the per-database launches': a tile is the same code on the same rows.
constexpr int kFusedSegs = (int)kFusedMaxSegments;
constexpr int kFusedGR = 4;
constexpr int kFusedQR = 4;

struct FusedSeg {
  const uint64_t *db;
  uint64_t n_l, groups;
  uint32_t P, n, m, sblocks;
  uint32_t qn, qr, qblocks;  // queries of the segment, per block (1..4), blocks
  uint32_t ptr0;             // its query pointers at ptrs[ptr0 ..], outputs after them
  uint32_t begin[8];         // first tile of the segment in each XCD's list
};

struct FusedArgs {
  ModSet ms;
  uint32_t *counters;        // [8], zero at launch: next tile of each XCD's list
  const uint64_t *ptrs;      // query and output pointers of every segment
  uint32_t nseg;
  uint32_t tiles[8];         // tiles in each XCD's list
  FusedSeg seg[kFusedSegs];
};

// a wave-uniform 64-bit value in scalar registers
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

template <int GR, int QR, int W, int NQMAX>
__device__ __forceinline__ void fused_tile(const FusedArgs &a, const FusedSeg &g, uint64_t cg, uint32_t qb,
                                           QPair (&lds)[2][NQMAX][64]) {
  const TileLevel lv{g.db, g.n_l, g.groups, g.P, g.n, g.m, g.sblocks, g.qn};
  const uint32_t q0 = qb * QR;
  const uint64_t *qbase[QR];
  uint64_t *obase[QR];
#pragma unroll
  for (int i = 0; i < QR; i++) {
    const uint32_t qi = g.ptr0 + min(q0 + i, g.qn - 1);
    qbase[i] = reinterpret_cast<const uint64_t *>(uniform64(a.ptrs[qi]));
    obase[i] = reinterpret_cast<uint64_t *>(uniform64(a.ptrs[qi + g.qn]));
  }
  batch3_tile<GR, QR, W, false>(a.ms, lv, qbase, obase, g.qn - q0, cg, lds);
}

template <int W>
__global__ void __launch_bounds__(64 * W) mac_fused_kernel(FusedArgs a) {
  constexpr int GR = kFusedGR;
  __shared__ QPair lds[2][kB3U * kFusedQR * 2][64];
  __shared__ uint32_t slot;
  const uint32_t home = blockIdx.x & 7;
  for (uint32_t k = 0; k < 8; k++) {
    const uint32_t x = (home + k) & 7;
    const uint32_t ntiles = a.tiles[x];
    for (;;) {
      if (threadIdx.x == 0) slot = atomicAdd(&a.counters[x], 1u);
      __syncthreads();
      const uint32_t t = __builtin_amdgcn_readfirstlane(slot);
      __syncthreads();  // every wave has read the slot before the next fetch writes it
      if (t >= ntiles) break;
      uint32_t s = 0;
      while (s + 1 < a.nseg && a.seg[s + 1].begin[x] <= t) s++;
      const FusedSeg &g = a.seg[s];
      const uint32_t j = t - g.begin[x];
      const uint64_t cg = x + 8ull * (j / g.qblocks);
      const uint32_t qb = j % g.qblocks;
      switch (g.qr) {
        case 4: fused_tile<GR, 4, W>(a, g, cg, qb, lds); break;
        case 3: fused_tile<GR, 3, W>(a, g, cg, qb, lds); break;
        case 2: fused_tile<GR, 2, W>(a, g, cg, qb, lds); break;
        default: fused_tile<GR, 1, W>(a, g, cg, qb, lds); break;
      }
    }
  }
}

// ------------------------------------------------------------------------------------
// Exact chains split along the element axis.  With few chains (a single query: 8192
// chains of 65536 steps) the element axis has to be split.  Write acc = rho + j*p with
// rho = acc mod p.  A step with lazy product r = rc + b*p (rc < p, b in {0,1}) gives
//     rho' = (rho + rc) mod p,   carry c = [rho + rc >= p],
//     j'   = max(j + b + c - 1, 0)
// (acc + r >= p  <=>  j + b + c >= 1).  The rho chain is a plain modular sum; the j chain
// is a max-plus map j -> max(j + D, M) whose composition is associative:
//     (D1,M1) then (D2,M2) = (D1 + D2, max(M1 + D2, M2)),
// so the final word rho_final + j_final * p is identical to the sequential chain's.

constexpr int32_t kNegInf = -(1 << 30);
constexpr int FIX_PARTS = 8;  // threads per chain in the scan kernel

__device__ __forceinline__ uint64_t addmod_c(uint64_t a, uint64_t b, uint64_t p) {
  uint64_t s = a + b;
  return s >= p ? s - p : s;
}

// ---- streaming MAC: per-block chain summaries, no inter-workgroup communication ----
//
// Every wave runs one (column, block of E consecutive elements) and keeps, per chain, the
// summary of its block's lazy products r_1..r_S read as a chain started from zero:
//   rho0 = A_S mod p,  J_t = t - floor(A_t / p),  Jend = J_S,  Jmax = max_t J_t,
//   Rstar = min { A_t mod p : J_t = Jmax },        A_t = r_1 + ... + r_t.
// Entering the block with residue rho_in (the chain's residue before it), the number of
// p's subtracted by step t is W_t = floor((rho_in + A_t) / p) = t - J_t + [rho_in + (A_t
// mod p) >= p], so the block's max-plus map is
//   D = [rho_in + rho0 >= p] - Jend,   M = D + Jmax - [rho_in + Rstar >= p]
// (the largest J_t - [rho_in + (A_t mod p) >= p] is Jmax unless every t reaching Jmax
// wraps, i.e. unless the smallest such residue does).  A second kernel scans the blocks
// of each chain in element order: residues by a sum mod p, maps by max-plus composition.
// Each DB and query word is read once; blocks never wait for one another.

struct StreamArgs {
  ModSet ms;
  const uint64_t *db;
  const uint64_t *q;
  uint64_t *out;
  uint64_t n_l;
  uint32_t columns, cpg, P, n, m, sblocks, log_sb;
  uint32_t E, nblk;     // elements per block (a multiple of the chunk), blocks per column
  uint64_t *rho0;       // [nblk][columns][2][64]
  uint64_t *rstar;      // same layout
  int2 *jj;             // (Jend, Jmax), same layout
};

// Block summary of one chain read from zero: Z_t = t*p - A_t, A_t the prefix sum of its
// lazy products (so Z_t = J_t*p - R_t with J_t = t - floor(A_t/p), R_t = A_t mod p), as a
// 96-bit two's-complement word (low 64 bits z, signed high word zh), and its running
// maximum (m, mh) over t = 1..E.  max_t Z_t is reached exactly where J_t is largest and,
// among those t, R_t smallest, so it carries (Jmax, Rstar); a step is one 96-bit add and
// one 96-bit compare, with no reduction modulo p.
struct ChainSum {
  uint64_t z, m;
  int32_t zh, mh;
};

// one element: Z += p - r (r < 2p), then the running maximum
__device__ __forceinline__ void zsum_step(ChainSum &s, uint64_t r, uint64_t p) {
  uint64_t u, v;
  const bool c = __builtin_add_overflow(s.z, p, &u);
  const bool b = __builtin_sub_overflow(u, r, &v);
  s.zh += (int32_t)c - (int32_t)b;
  s.z = v;
  const bool take = s.zh > s.mh || (s.zh == s.mh && s.z > s.m);
  s.mh = take ? s.zh : s.mh;
  s.m = take ? s.z : s.m;
}

// Z += d for a signed 64-bit d (96-bit add: low word with carry, high word plus the carry
// and d's sign extension)
__device__ __forceinline__ void add96(uint64_t z, int32_t zh, int64_t d, uint64_t &oz, int32_t &ozh) {
  uint64_t u;
  const bool c = __builtin_add_overflow(z, (uint64_t)d, &u);
  ozh = zh + (int32_t)c + (d < 0 ? -1 : 0);
  oz = u;
}

// four elements at once: with s_k = sum_{j<=k} (p - r_j) (|s_k| <= 4p < 2^63), the four
// new Z values are Z + s_k, so their maximum is Z + max_k s_k: one 96-bit compare per
// four steps instead of four
__device__ __forceinline__ void zsum_group4(ChainSum &s, uint64_t r0, uint64_t r1, uint64_t r2, uint64_t r3,
                                            uint64_t pv) {
  const int64_t s1 = (int64_t)(pv - r0);
  const int64_t s2 = s1 + (int64_t)(pv - r1);
  const int64_t s3 = s2 + (int64_t)(pv - r2);
  const int64_t s4 = s3 + (int64_t)(pv - r3);
  const int64_t mx = max(max(s1, s2), max(s3, s4));
  uint64_t cl;
  int32_t ch;
  add96(s.z, s.zh, mx, cl, ch);
  const bool take = ch > s.mh || (ch == s.mh && cl > s.m);
  s.mh = take ? ch : s.mh;
  s.m = take ? cl : s.m;
  add96(s.z, s.zh, s4, s.z, s.zh);
}

// J = ceil(Z/p) and R = J*p - Z in [0, p), for |Z| < 2^81 (blocks of at most 2^20
// elements): N = -Z + 2^21*p >= 0 is split as H*2^61 + L = H*p + (H*delta + L), and
// H*delta + L < 2p, so floor(N/p) = H + [H*delta + L >= p].
__device__ __forceinline__ void z_to_j_r(uint64_t z, int32_t zh, uint64_t p, uint32_t delta, int32_t &J,
                                         uint64_t &R) {
  const __int128 Z = (__int128)(((unsigned __int128)(__int128)zh << 64) | z);
  const unsigned __int128 N = (unsigned __int128)(((__int128)p << 21) - Z);
  const uint64_t H = (uint64_t)(N >> 61), L = (uint64_t)N & ((1ULL << 61) - 1);
  const uint64_t S = H * delta + L;
  const uint64_t b = S >= p ? 1 : 0;
  J = (int32_t)((1 << 21) - (int64_t)(H + b));
  R = S - b * p;
}

// Loads of one chunk of U elements (x, a, b words of this lane): xp / qp are the chunk's
// wave-uniform row bases and lo the lane's byte offset in a row (scalar base plus one
// 32-bit vector offset per load, no per-lane address arithmetic); the elements at or
// past `valid` read as zero and are not summed.
template <int U>
__device__ __forceinline__ void stream_load(const uint64_t *xp, const uint64_t *qp, uint32_t lo, uint64_t xstride,
                                            uint64_t qstride, uint64_t mn, int valid, uint64_t (&xv)[U],
                                            uint64_t (&av)[U], uint64_t (&bv)[U]) {
  if (valid >= U) {
#pragma unroll
    for (int j = 0; j < U; j++) {
      xv[j] = __builtin_nontemporal_load(lane_word(xp + j * xstride, lo));
      av[j] = __builtin_nontemporal_load(lane_word(qp + j * qstride, lo));
      bv[j] = __builtin_nontemporal_load(lane_word(qp + j * qstride + mn, lo));
    }
  } else {
#pragma unroll
    for (int j = 0; j < U; j++) {
      xv[j] = j < valid ? *lane_word(xp + j * xstride, lo) : 0;
      av[j] = j < valid ? *lane_word(qp + j * qstride, lo) : 0;
      bv[j] = j < valid ? *lane_word(qp + j * qstride + mn, lo) : 0;
    }
  }
}

template <int U>
__device__ __forceinline__ void stream_sum(ChainSum &sa, ChainSum &sb, int valid, const uint64_t (&xv)[U],
                                           const uint64_t (&av)[U], const uint64_t (&bv)[U], uint64_t p,
                                           uint64_t pv, uint32_t delta) {
  if (valid >= U) {
    static_assert(U % 4 == 0, "chunks are whole groups of four elements");
#pragma unroll
    for (int j = 0; j < U; j += 4) {
      uint64_t ra[4], rb[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        ra[i] = shoup_lazy_asm(xv[j + i], av[j + i], shoup_quot(av[j + i], p, delta), delta);
        rb[i] = shoup_lazy_asm(xv[j + i], bv[j + i], shoup_quot(bv[j + i], p, delta), delta);
      }
      zsum_group4(sa, ra[0], ra[1], ra[2], ra[3], pv);
      zsum_group4(sb, rb[0], rb[1], rb[2], rb[3], pv);
    }
  } else {
#pragma unroll
    for (int j = 0; j < U; j++) {
      if (j < valid) {
        zsum_step(sa, shoup_lazy_asm(xv[j], av[j], shoup_quot(av[j], p, delta), delta), p);
        zsum_step(sb, shoup_lazy_asm(xv[j], bv[j], shoup_quot(bv[j], p, delta), delta), p);
      }
    }
  }
}

// One wave = one (column, block), read in chunks of U elements.  PF 0: a chunk's loads,
// then its arithmetic (other waves fill the wait).  PF 1: two register sets in turn, so
// the next chunk's loads are in flight while the current chunk's products and chain
// steps run.  PF 2 (rolling): one register set; as each group of four elements' products
// free their registers, the next chunk's loads for those four elements go out, so a wave
// keeps about a chunk of loads in flight through its arithmetic at PF 0's register cost.
template <int U, int PF>
__global__ void __launch_bounds__(256) mac_stream_kernel(StreamArgs a) {
  const uint32_t lane = threadIdx.x & 63;
  // wave-uniform: global wave -> (block, column); adjacent waves take adjacent columns
  const uint32_t gw = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  const uint32_t blk = gw / a.columns, column = gw - blk * a.columns;
  if (blk >= a.nblk) return;
  const uint32_t g = column / a.cpg, cw = column - g * a.cpg;
  const uint32_t c = (cw >> a.log_sb) % a.m, sb = cw & (a.sblocks - 1);
  const uint64_t p = a.ms.p[c];
  uint64_t pv = p;
  asm volatile("" : "+v"(pv));  // the modulus in VGPRs too, for the p - r of every step
  const uint32_t delta = a.ms.delta[c];
  const uint64_t mn = (uint64_t)a.m * a.n;
  const uint64_t xstride = (uint64_t)a.P * mn, qstride = 2 * mn;
  const uint64_t e0 = (uint64_t)blk * a.E;
  const int64_t n = (int64_t)(min<uint64_t>(a.n_l, e0 + a.E) - e0);  // elements in this block
  // wave-uniform row bases; the lane reads word `lane` of each 64-word row segment
  const uint64_t *xp = a.db + ((uint64_t)g * a.n_l + e0) * xstride + (uint64_t)cw * 64;
  const uint64_t *qp = a.q + e0 * qstride + (uint64_t)c * a.n + sb * 64;
  const uint32_t lo = lane * 8;
  ChainSum sa{0, 0, 0, INT32_MIN}, sbs{0, 0, 0, INT32_MIN};
  if (PF == 0) {
    // one register set: the chunk's loads, then its arithmetic (other waves fill the gap)
    for (int64_t e = 0; e < n; e += U) {
      uint64_t xv[U], av[U], bv[U];
      const int valid = (int)min<int64_t>(n - e, U);
      stream_load<U>(xp + e * xstride, qp + e * qstride, lo, xstride, qstride, mn, valid, xv, av, bv);
      stream_sum<U>(sa, sbs, valid, xv, av, bv, p, pv, delta);
    }
  } else if (PF == 2) {
    uint64_t xv[U], av[U], bv[U];
    stream_load<U>(xp, qp, lo, xstride, qstride, mn, (int)min<int64_t>(n, U), xv, av, bv);
    int64_t e = 0;
    // the registers hold chunk e (loaded or in flight); while it and the next are whole,
    // each group of four elements: products, then the next chunk's loads into the freed
    // registers, then the group's summary step
    for (; e + 2 * U <= n; e += U) {
      const uint64_t *xn = xp + (e + U) * xstride, *qn = qp + (e + U) * qstride;
#pragma unroll
      for (int j = 0; j < U; j += 4) {
        uint64_t ra[4], rb[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          ra[i] = shoup_lazy_asm(xv[j + i], av[j + i], shoup_quot(av[j + i], p, delta), delta);
          rb[i] = shoup_lazy_asm(xv[j + i], bv[j + i], shoup_quot(bv[j + i], p, delta), delta);
        }
        __builtin_amdgcn_sched_barrier(0);  // the loads go out here, not at their first use
#pragma unroll
        for (int i = 0; i < 4; i++) {
          xv[j + i] = __builtin_nontemporal_load(lane_word(xn + (j + i) * xstride, lo));
          av[j + i] = __builtin_nontemporal_load(lane_word(qn + (j + i) * qstride, lo));
          bv[j + i] = __builtin_nontemporal_load(lane_word(qn + (j + i) * qstride + mn, lo));
        }
        __builtin_amdgcn_sched_barrier(0);
        zsum_group4(sa, ra[0], ra[1], ra[2], ra[3], pv);
        zsum_group4(sbs, rb[0], rb[1], rb[2], rb[3], pv);
      }
    }
    // the last one or two chunks (the current whole or partial, the next partial or none)
    for (; e < n; e += U) {
      stream_sum<U>(sa, sbs, (int)min<int64_t>(n - e, U), xv, av, bv, p, pv, delta);
      if (e + U < n)
        stream_load<U>(xp + (e + U) * xstride, qp + (e + U) * qstride, lo, xstride, qstride, mn,
                       (int)min<int64_t>(n - e - U, U), xv, av, bv);
    }
  } else {
  uint64_t x0[U], a0[U], b0[U], x1[U], a1[U], b1[U];
  stream_load<U>(xp, qp, lo, xstride, qstride, mn, (int)min<int64_t>(n, U), x0, a0, b0);
  for (int64_t e = 0; e < n; e += 2 * U) {
    if (e + U < n)
      stream_load<U>(xp + (e + U) * xstride, qp + (e + U) * qstride, lo, xstride, qstride, mn,
                     (int)min<int64_t>(n - e - U, U), x1, a1, b1);
    stream_sum<U>(sa, sbs, (int)min<int64_t>(n - e, U), x0, a0, b0, p, pv, delta);
    if (e + U >= n) break;
    if (e + 2 * U < n)
      stream_load<U>(xp + (e + 2 * U) * xstride, qp + (e + 2 * U) * qstride, lo, xstride, qstride, mn,
                     (int)min<int64_t>(n - e - 2 * U, U), x0, a0, b0);
    stream_sum<U>(sa, sbs, (int)min<int64_t>(n - e - U, U), x1, a1, b1, p, pv, delta);
  }
  }
  // (rho0, Jend) from Z at the block end, (Rstar, Jmax) from its maximum
  int32_t ja, jma, jb, jmb;
  uint64_t ra, rsa, rb, rsb;
  z_to_j_r(sa.z, sa.zh, p, delta, ja, ra);
  z_to_j_r(sa.m, sa.mh, p, delta, jma, rsa);
  z_to_j_r(sbs.z, sbs.zh, p, delta, jb, rb);
  z_to_j_r(sbs.m, sbs.mh, p, delta, jmb, rsb);
  const uint64_t i = ((uint64_t)gw * 2) * 64 + lane;  // gw = blk * columns + column
  a.rho0[i] = ra;
  a.rho0[i + 64] = rb;
  a.rstar[i] = rsa;
  a.rstar[i + 64] = rsb;
  a.jj[i] = make_int2(ja, jma);
  a.jj[i + 64] = make_int2(jb, jmb);
}

// Scan the block summaries of every chain in element order and emit the reply words.
// Workgroup = one column: 128 chains x FIX_PARTS threads, each over a contiguous run of
// blocks.  Pass 1 sums each run's residues; pass 2 evaluates the run's maps from the now
// known incoming residues; the runs' maps are then composed in order.
__global__ void __launch_bounds__(128 * FIX_PARTS) mac_stream_fixup_kernel(StreamArgs a) {
  __shared__ uint64_t part_rho[FIX_PARTS][128];
  __shared__ int2 part_map[FIX_PARTS][128];
  const uint32_t column = blockIdx.x;
  const uint32_t cl = threadIdx.x & 127, pi = threadIdx.x >> 7;
  const uint32_t r = column >> a.log_sb;
  const uint32_t c = r % a.m;
  const uint64_t p = a.ms.p[c];
  const uint32_t per = (a.nblk + FIX_PARTS - 1) / FIX_PARTS;
  const uint32_t t0 = min(a.nblk, pi * per), t1 = min(a.nblk, t0 + per);
  auto idx = [&](uint32_t t) { return ((uint64_t)t * a.columns + column) * 128 + cl; };
  // a run of up to KB blocks (the usual case: about 64 blocks per column) is loaded into
  // registers at once, so its loads are in flight together instead of one per iteration
  constexpr uint32_t KB = 8;
  const bool cached = t1 - t0 <= KB;
  uint64_t r0v[KB], rsv[KB];
  int2 jv[KB];
  uint64_t s = 0;
  if (cached) {
#pragma unroll
    for (uint32_t i = 0; i < KB; i++)
      if (t0 + i < t1) {
        r0v[i] = a.rho0[idx(t0 + i)];
        rsv[i] = a.rstar[idx(t0 + i)];
        jv[i] = a.jj[idx(t0 + i)];
      }
#pragma unroll
    for (uint32_t i = 0; i < KB; i++)
      if (t0 + i < t1) s = addmod_c(s, r0v[i], p);
  } else {
    for (uint32_t t = t0; t < t1; t++) s = addmod_c(s, a.rho0[idx(t)], p);
  }
  part_rho[pi][cl] = s;
  __syncthreads();
  uint64_t rho_in = 0;
  for (uint32_t i = 0; i < pi; i++) rho_in = addmod_c(rho_in, part_rho[i][cl], p);
  int2 f = make_int2(0, kNegInf);
  auto step = [&](uint64_t r0, uint64_t rs, int2 j) {
    const int32_t D = (rho_in + r0 >= p ? 1 : 0) - j.x;
    const int32_t M = D + j.y - (rho_in + rs >= p ? 1 : 0);
    f = make_int2(f.x + D, max(f.y + D, M));
    rho_in = addmod_c(rho_in, r0, p);
  };
  if (cached) {
#pragma unroll
    for (uint32_t i = 0; i < KB; i++)
      if (t0 + i < t1) step(r0v[i], rsv[i], jv[i]);
  } else {
    for (uint32_t t = t0; t < t1; t++) step(a.rho0[idx(t)], a.rstar[idx(t)], a.jj[idx(t)]);
  }
  part_map[pi][cl] = f;
  __syncthreads();
  if (pi != FIX_PARTS - 1) return;
  int32_t jfin = 0;
#pragma unroll
  for (int i = 0; i < FIX_PARTS; i++) jfin = max(jfin + part_map[i][cl].x, part_map[i][cl].y);
  const uint32_t chain = cl >> 6, lane = cl & 63;
  const uint32_t sb = column & (a.sblocks - 1);
  const uint64_t gk = r / a.m;
  const uint64_t mn = (uint64_t)a.m * a.n;
  a.out[gk * 2 * mn + chain * mn + (uint64_t)c * a.n + sb * 64 + lane] = rho_in + (uint64_t)jfin * p;
}

}  // namespace

namespace {
// Streaming launch: block summaries into scratch, then the per-column scan.
template <int U, int PF>
int launch_stream(xpg_ctx *ctx, StreamArgs &st) {
  const uint64_t recs = (uint64_t)st.nblk * st.columns * 128;
  const uint64_t waves = (uint64_t)st.nblk * st.columns;
  if (waves > 0xffffffffULL - 3) return set_err(ctx, XPG_EINVAL, "MAC grid too large");
  void *scr = nullptr;
  int rc = get_scratch(ctx, recs * 24, &scr);
  if (rc) return rc;
  st.rho0 = (uint64_t *)scr;
  st.rstar = st.rho0 + recs;
  st.jj = (int2 *)(st.rstar + recs);
  hipEvent_t k0 = nullptr, k1 = nullptr;
  timing_begin(ctx, &k0, &k1, 1);
  hipLaunchKernelGGL((mac_stream_kernel<U, PF>), dim3((uint32_t)((waves + 3) / 4)), dim3(256), 0, ctx->stream, st);
  timing_end(ctx, k0, k1, 1);
  if (hipGetLastError() != hipSuccess) return set_err(ctx, XPG_EDEVICE, "MAC launch");
  hipLaunchKernelGGL(mac_stream_fixup_kernel, dim3(st.columns), dim3(128 * FIX_PARTS), 0, ctx->stream, st);
  if (hipGetLastError() != hipSuccess) return set_err(ctx, XPG_EDEVICE, "MAC fixup launch");
  return XPG_OK;
}

uint64_t columns_of(const xpg_params *prm, uint64_t groups, uint32_t P) {
  return groups * P * (uint64_t)prm->m * (prm->n / 64);
}

// Elements per block that give about four waves per SIMD, all resident at once: one
// pass over the level, and few block summaries (their writes interleave with the
// level's reads).
uint64_t auto_e(int cu_count, uint64_t n_l, uint64_t columns, int U) {
  const uint64_t target = 16ull * (uint64_t)cu_count;
  const uint64_t nb = std::max<uint64_t>(1, target / std::max<uint64_t>(1, columns));
  const uint64_t E = (n_l + nb - 1) / nb;
  return (E + U - 1) / U * U;
}

// Elements per streaming block: E as asked (0 = auto_e), at most 2^20 (the block
// summary's 96-bit word holds 2^20 steps), a whole number of chunks.
uint64_t stream_block_elems(int cu_count, uint64_t n_l, uint64_t columns, int U, uint64_t E) {
  if (!E) E = auto_e(cu_count, n_l, columns, U);
  E = std::min<uint64_t>(E, 1u << 20);
  return (E + U - 1) / U * U;
}

// The streaming MAC for one level: U elements per load chunk, pf the load schedule (0
// load then compute, 1 double-buffered chunks, 2 rolling), E elements per block (0 =
// auto_e).
int launch_stream_level(xpg_ctx *ctx, const MacArgs &a, uint32_t n, uint64_t groups, int U, int pf, uint64_t E) {
  StreamArgs st{};
  st.ms = a.ms;
  st.db = a.db;
  st.q = a.q;
  st.out = a.out;
  st.n_l = a.n_l;
  st.P = a.P;
  st.n = n;
  st.m = a.m;
  st.sblocks = n / 64;
  st.log_sb = (uint32_t)__builtin_ctz(st.sblocks);
  st.cpg = a.P * a.m * st.sblocks;
  if (groups * st.cpg > 0xffffffffULL) return set_err(ctx, XPG_EINVAL, "MAC grid too large");
  st.columns = (uint32_t)(groups * st.cpg);
  if (U != 4 && U != 8 && U != 16) return set_err(ctx, XPG_EINVAL, "streaming MAC chunk not built");
  if (pf < 0 || pf > 2) return set_err(ctx, XPG_EINVAL, "streaming MAC load schedule not built");
  E = stream_block_elems(ctx->cu_count, a.n_l, st.columns, U, E);
  if (E > 0xffffffffULL || (a.n_l + E - 1) / E > 0xffffffffULL) return set_err(ctx, XPG_EINVAL, "MAC grid too large");
  st.E = (uint32_t)E;
  st.nblk = (uint32_t)((a.n_l + E - 1) / E);
  switch (U * 4 + pf) {
    case 16: return launch_stream<4, 0>(ctx, st);
    case 17: return launch_stream<4, 1>(ctx, st);
    case 18: return launch_stream<4, 2>(ctx, st);
    case 64: return launch_stream<16, 0>(ctx, st);
    case 65: return launch_stream<16, 1>(ctx, st);
    case 66: return launch_stream<16, 2>(ctx, st);
    case 33: return launch_stream<8, 1>(ctx, st);
    case 34: return launch_stream<8, 2>(ctx, st);
    default: return launch_stream<8, 0>(ctx, st);
  }
}

}  // namespace

void batch_default_shape(uint32_t qn, uint64_t groups, int *gr, int *qr) {
  if (!*qr) *qr = qn >= 4 ? 4 : (qn >= 2 ? 2 : 1);
  if (!*gr) *gr = groups >= 8 && *qr <= 2 ? 8 : (groups >= 4 ? 4 : 1);
}

bool batch_shape_built(int GR, int QR) {
  const int key = GR * 10 + QR;
  return key == 11 || key == 12 || key == 14 || key == 41 || key == 42 || key == 44 || key == 81 || key == 82 ||
         key == 161;
}

bool batch_family_built(int family, int GR, int QR) {
  const int key = GR * 10 + QR;
  switch (family) {
    case 2: return batch_shape_built(GR, QR);
    case 3: return key == 41 || key == 42 || key == 44 || key == 81 || key == 82;
    default: return false;
  }
}

bool batch3_shape_built(int GR, int QR) { return batch_family_built(3, GR, QR); }


// Lanes at which one chain per lane already fills the GPU (256 CUs x 32 wavefronts).
static const uint64_t kDirectLanes = 256ull * 32 * 64;

uint64_t stream_auto_e(const xpg_ctx *ctx, const xpg_params *prm, uint64_t n_l, uint64_t groups, uint32_t P, int u) {
  return auto_e(ctx->cu_count, n_l, columns_of(prm, groups, P), u ? u : 8);
}

uint64_t level_scratch_bytes(const xpg_ctx *ctx, const xpg_params *prm, uint64_t n_l, uint64_t groups, uint32_t P,
                             const LevelCfg &cfg) {
  if (cfg.kernel != 3) return 0;
  const uint64_t columns = columns_of(prm, groups, P);
  const int U = cfg.a ? cfg.a : 8;
  const uint64_t E = stream_block_elems(ctx->cu_count, n_l, columns, U, (uint64_t)cfg.c);
  return ((n_l + E - 1) / E) * columns * 128 * 24;
}

LevelCfg resolve_level_cfg(const xpg_ctx *ctx, const xpg_params *prm, uint64_t n_l, uint64_t groups, uint32_t P) {
  LevelCfg cfg;
  const TuneKey key{0, prm->n, prm->m, P, n_l, groups, 0};
  if (options_pinned(ctx) || !tune_lookup(ctx->device, key, &cfg)) cfg = default_level_cfg(ctx, prm, n_l, groups, P);
  return cfg;
}

LevelCfg default_level_cfg(const xpg_ctx *ctx, const xpg_params *prm, uint64_t n_l, uint64_t groups, uint32_t P) {
  const uint64_t lanes = groups * P * (uint64_t)prm->m * prm->n;
  // the streaming kernel's columns are 64 slots wide: degrees below 64 always take one
  // chain per lane
  const bool one_chain_per_lane =
      prm->n < 64 || ctx->mac_mode == 1 || (ctx->mac_mode == 0 && (lanes >= kDirectLanes || n_l <= 64));
  LevelCfg c;
  if (one_chain_per_lane && groups >= 4 && ctx->mac_mode != 1) {
    // many groups share each query word: register-blocked over groups
    c.kernel = 4;
    c.b = 1;
  } else if (one_chain_per_lane) {
    c.kernel = 1;
  } else {
    c.kernel = 3;
    c.a = ctx->stream_u ? ctx->stream_u : 8;
    // load schedule: option stream_pf 0 (unset) = the rule's rolling loads, 3 = load then
    // compute, 1 / 2 as they are (LevelCfg.b: 0 load then compute, 1 double-buffered, 2 rolling)
    c.b = ctx->stream_pf == 0 ? 2 : ctx->stream_pf == 3 ? 0 : ctx->stream_pf;
    // elements per block: the option, else for the rolling loads twice the one-pass size
    // (two waves per SIMD keep enough loads in flight, with longer runs per wave)
    c.c = ctx->stream_e ? ctx->stream_e
                        : c.b == 2 ? (int)std::min<uint64_t>(1u << 20, 2 * stream_auto_e(ctx, prm, n_l, groups, P, c.a))
                                   : 0;
  }
  return c;
}

int launch_mac_level_cfg(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *db, const uint64_t *q,
                         uint64_t n_l, uint64_t groups, uint32_t P, uint64_t *out, const LevelCfg &cfg) {
  if (cfg.kernel == 4) {
    uint64_t *outp = out;
    return launch_mac_batch_cfg(ctx, prm, db, &q, &outp, 1, n_l, groups, P, cfg.a, 1, cfg.c);
  }
  if (cfg.kernel != 1 && cfg.kernel != 3) return set_err(ctx, XPG_EINVAL, "unknown MAC kernel");
  if (cfg.kernel == 3 && prm->n < 64) return set_err(ctx, XPG_EINVAL, "MAC kernel needs n >= 64");
  MacArgs a;
  a.ms = modset_of(prm);
  a.db = db;
  a.q = q;
  a.out = out;
  a.n_l = n_l;
  a.groups = groups;
  a.P = P;
  a.n = prm->n;
  a.m = prm->m;
  a.logn = prm->logn;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  timing_begin(ctx, &e0, &e1);
  int rc = XPG_OK;
  if (cfg.kernel == 3) {
    rc = launch_stream_level(ctx, a, prm->n, groups, cfg.a ? cfg.a : 8, cfg.b, (uint64_t)cfg.c);
  } else if (cfg.kernel == 1) {
    const uint64_t lanes = groups * P * (uint64_t)prm->m * prm->n;
    const uint64_t blocks = (lanes + 255) / 256;
    if (blocks > 0x7fffffffULL) return set_err(ctx, XPG_EINVAL, "MAC grid too large");
    hipEvent_t k0 = nullptr, k1 = nullptr;
    timing_begin(ctx, &k0, &k1, 1);
    hipLaunchKernelGGL(mac_direct_kernel, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, a);
    timing_end(ctx, k0, k1, 1);
    if (hipGetLastError() != hipSuccess) rc = set_err(ctx, XPG_EDEVICE, "MAC launch");
  }
  timing_end(ctx, e0, e1);
  return rc;
}

// Explicit options (a kernel or a tile shape asked for) win over the tuning table.
bool options_pinned(const xpg_ctx *ctx) {
  return ctx->mac_mode || ctx->stream_e || ctx->stream_u || ctx->stream_pf;
}

int launch_mac_level(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *db, const uint64_t *q,
                     uint64_t n_l, uint64_t groups, uint32_t P, uint64_t *out) {
  return launch_mac_level_cfg(ctx, prm, db, q, n_l, groups, P, out, resolve_level_cfg(ctx, prm, n_l, groups, P));
}

int launch_mac_batch(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *db, const uint64_t *const *q,
                     uint64_t *const *out, uint32_t qn, uint64_t n_l, uint64_t groups, uint32_t P) {
  LevelCfg cfg;
  const TuneKey key{1, prm->n, prm->m, P, n_l, groups, tune_qn_bucket(qn)};
  if (options_pinned(ctx) || !tune_lookup(ctx->device, key, &cfg) || cfg.kernel != 4) cfg = LevelCfg{};
  return launch_mac_batch_cfg(ctx, prm, db, q, out, qn, n_l, groups, P, cfg.a, cfg.b, cfg.c);
}

int launch_mac_batch_cfg(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *db, const uint64_t *const *q,
                         uint64_t *const *out, uint32_t qn, uint64_t n_l, uint64_t groups, uint32_t P, int gr,
                         int qr, int family) {
  if (qn == 0 || qn > (uint32_t)kMaxBatch) return set_err(ctx, XPG_EINVAL, "batch of 1..64 queries per launch");
  if (prm->n < 64) return set_err(ctx, XPG_EINVAL, "MAC kernel needs n >= 64");
  // register blocks: GR rows x QR queries per lane (0 = by shape); family 3 shares the
  // query words of a workgroup's four waves through LDS, family 2 loads them per wave
  // (0 = by shape)
  int QR = qr, GR = gr;
  batch_default_shape(qn, groups, &GR, &QR);
  if (family == 0) family = batch3_shape_built(GR, QR) ? 3 : 2;
  const int key = GR * 10 + QR;
  if (!batch_family_built(family, GR, QR)) return set_err(ctx, XPG_EINVAL, "register block shape not built");
  // a last query block of fewer than QR queries would compute padding queries: those r
  // queries take a launch of their own, with the shape the fixed rules give r queries
  const uint32_t rest = qn % (uint32_t)QR;
  if (qn > (uint32_t)QR && rest) {
    const int rc = launch_mac_batch_cfg(ctx, prm, db, q, out, qn - rest, n_l, groups, P, GR, QR, family);
    if (rc) return rc;
    return launch_mac_batch_cfg(ctx, prm, db, q + (qn - rest), out + (qn - rest), rest, n_l, groups, P, 0, 0, 0);
  }
  Batch2Args b2{};
  b2.ms = modset_of(prm);
  b2.db = db;
  for (uint32_t i = 0; i < qn; i++) {
    b2.q[i] = q[i];
    b2.out[i] = out[i];
  }
  b2.n_l = n_l;
  b2.groups = groups;
  b2.P = P;
  b2.n = prm->n;
  b2.m = prm->m;
  b2.sblocks = prm->n / 64;
  b2.qn = qn;
  b2.qblocks = (qn + QR - 1) / QR;
  // family 2: a block = GR groups x P polys of one column; family 3: W*GR rows
  constexpr int W = 4;
  b2.colgroups = family == 3 ? ((groups * P + W * GR - 1) / (W * GR)) * prm->m * b2.sblocks
                             : ((groups + GR - 1) / GR) * P * prm->m * b2.sblocks;
  const uint64_t blocks = ((b2.colgroups + 7) / 8) * 8 * b2.qblocks;
  if (blocks > 0x7fffffffULL) return set_err(ctx, XPG_EINVAL, "batched MAC grid too large");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  timing_begin(ctx, &e0, &e1);
  hipEvent_t k0 = nullptr, k1 = nullptr;
  timing_begin(ctx, &k0, &k1, 2);
  if (ctx->timing && !ctx->ws) ctx->batch_steps += (uint64_t)qn * groups * P * prm->m * prm->n * 2 * n_l;
  const dim3 wg3(64 * W), grid((uint32_t)blocks);
  if (family == 3) {
    switch (key) {
      case 82: hipLaunchKernelGGL((mac_batch3_kernel<8, 2, 4>), grid, wg3, 0, ctx->stream, b2); break;
      case 81: hipLaunchKernelGGL((mac_batch3_kernel<8, 1, 4>), grid, wg3, 0, ctx->stream, b2); break;
      case 44: hipLaunchKernelGGL((mac_batch3_kernel<4, 4, 4>), grid, wg3, 0, ctx->stream, b2); break;
      case 42: hipLaunchKernelGGL((mac_batch3_kernel<4, 2, 4>), grid, wg3, 0, ctx->stream, b2); break;
      default: hipLaunchKernelGGL((mac_batch3_kernel<4, 1, 4>), grid, wg3, 0, ctx->stream, b2); break;
    }
  } else switch (key) {
    case 161: hipLaunchKernelGGL((mac_batch2_kernel<16, 1>), dim3((uint32_t)blocks), dim3(64), 0, ctx->stream, b2); break;
    case 82: hipLaunchKernelGGL((mac_batch2_kernel<8, 2>), dim3((uint32_t)blocks), dim3(64), 0, ctx->stream, b2); break;
    case 81: hipLaunchKernelGGL((mac_batch2_kernel<8, 1>), dim3((uint32_t)blocks), dim3(64), 0, ctx->stream, b2); break;
    case 44: hipLaunchKernelGGL((mac_batch2_kernel<4, 4>), dim3((uint32_t)blocks), dim3(64), 0, ctx->stream, b2); break;
    case 42: hipLaunchKernelGGL((mac_batch2_kernel<4, 2>), dim3((uint32_t)blocks), dim3(64), 0, ctx->stream, b2); break;
    case 41: hipLaunchKernelGGL((mac_batch2_kernel<4, 1>), dim3((uint32_t)blocks), dim3(64), 0, ctx->stream, b2); break;
    case 14: hipLaunchKernelGGL((mac_batch2_kernel<1, 4>), dim3((uint32_t)blocks), dim3(64), 0, ctx->stream, b2); break;
    case 12: hipLaunchKernelGGL((mac_batch2_kernel<1, 2>), dim3((uint32_t)blocks), dim3(64), 0, ctx->stream, b2); break;
    default: hipLaunchKernelGGL((mac_batch2_kernel<1, 1>), dim3((uint32_t)blocks), dim3(64), 0, ctx->stream, b2); break;
  }
  const int rc = hipGetLastError() != hipSuccess ? set_err(ctx, XPG_EDEVICE, "batched MAC launch") : XPG_OK;
  timing_end(ctx, k0, k1, 2);
  timing_end(ctx, e0, e1);
  return rc;
}

// The last L steps of the level-0 chains of qn (1..64) queries whose outputs already hold
// the residues of their first elements (the matrix-core level 0): tail_db holds each
// group's last L elements (groups*L rows x P x m x n), q[i] points at query i's first
// tail element.  Register blocks of 4 rows x QR queries with the query words shared
// through LDS (mac_batch3_kernel in its TAIL form).
int launch_mac_tail(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *tail_db, const uint64_t *const *q,
                    uint64_t *const *out, uint32_t qn, uint64_t L, uint64_t groups, uint32_t P, const TailStage &ts,
                    uint32_t qfirst) {
  if (qn == 0 || qn > (uint32_t)kMaxBatch) return set_err(ctx, XPG_EINVAL, "batch of 1..64 queries per launch");
  if (prm->n < 64) return set_err(ctx, XPG_EINVAL, "MAC kernel needs n >= 64");
  constexpr int GR = 4, W = 4;
  const uint32_t QR = qn >= 4 ? 4 : qn >= 2 ? 2 : 1;
  const uint32_t rest = qn % QR;
  if (rest && qn > QR) {
    const int rc = launch_mac_tail(ctx, prm, tail_db, q, out, qn - rest, L, groups, P, ts, qfirst);
    if (rc) return rc;
    return launch_mac_tail(ctx, prm, tail_db, q + (qn - rest), out + (qn - rest), rest, L, groups, P, ts,
                           qfirst + (qn - rest));
  }
  Batch2Args b2{};
  b2.ms = modset_of(prm);
  b2.db = tail_db;
  for (uint32_t i = 0; i < qn; i++) {
    b2.q[i] = q[i];
    b2.out[i] = out[i];
  }
  b2.n_l = L;
  b2.groups = groups;
  b2.P = P;
  b2.n = prm->n;
  b2.m = prm->m;
  b2.sblocks = prm->n / 64;
  b2.qn = qn;
  b2.qblocks = (qn + QR - 1) / QR;
  b2.colgroups = ((groups * P + W * GR - 1) / (W * GR)) * prm->m * b2.sblocks;
  b2.stage = ts.stage;
  b2.st_cb = ts.cb;
  b2.st_rb = ts.rb;
  b2.st_mt = ts.mt;
  b2.qfirst = qfirst;
  const uint64_t blocks = ((b2.colgroups + 7) / 8) * 8 * b2.qblocks;
  if (blocks > 0x7fffffffULL) return set_err(ctx, XPG_EINVAL, "tail MAC grid too large");
  const dim3 wg(64 * W), grid((uint32_t)blocks);
  switch (QR) {
    case 4: hipLaunchKernelGGL((mac_batch3_kernel<GR, 4, W, true>), grid, wg, 0, ctx->stream, b2); break;
    case 2: hipLaunchKernelGGL((mac_batch3_kernel<GR, 2, W, true>), grid, wg, 0, ctx->stream, b2); break;
    default: hipLaunchKernelGGL((mac_batch3_kernel<GR, 1, W, true>), grid, wg, 0, ctx->stream, b2); break;
  }
  return hipGetLastError() != hipSuccess ? set_err(ctx, XPG_EDEVICE, "tail MAC launch") : XPG_OK;
}

size_t fused_table_bytes(uint32_t queries) { return 64 + (size_t)queries * 2 * sizeof(uint64_t); }

uint32_t fused_segments(uint32_t qn) { return qn > (uint32_t)kFusedQR && qn % kFusedQR ? 2u : 1u; }

// Workgroups of the fused kernel resident at once on the device (its persistent grid).
static uint32_t fused_grid(const xpg_ctx *ctx) {
  static const int per_cu = [] {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void *>(&mac_fused_kernel<4>), 256,
                                                     0) != hipSuccess || nb < 1)
      nb = 1;
    return nb;
  }();
  return (uint32_t)(ctx->cu_count * per_cu + 7) / 8 * 8;
}

int launch_mac_fused(xpg_ctx *ctx, const FusedPart *parts, uint32_t nparts, void *dev_table) {
  if (nparts == 0) return XPG_OK;
  if (!parts || !dev_table) return set_err(ctx, XPG_EINVAL, "fused launch without parts or table");
  constexpr int W = 4, GR = kFusedGR;
  const xpg_params *prm0 = parts[0].prm;
  // segments: a chunk's blocks of four queries, and its one block of the 1..3 left over
  struct SegH {
    FusedSeg s;
    const FusedPart *part;
    uint32_t q_first;
  };
  std::vector<SegH> segs;
  uint32_t queries = 0;
  uint64_t steps = 0;
  for (uint32_t pi = 0; pi < nparts; pi++) {
    const FusedPart &pt = parts[pi];
    const xpg_params *prm = pt.prm;
    if (pt.qn == 0 || pt.qn > (uint32_t)kMaxBatch) return set_err(ctx, XPG_EINVAL, "fused part of 1..64 queries");
    if (prm->n < 64 || prm->n % 64) return set_err(ctx, XPG_EINVAL, "MAC kernel needs n a multiple of 64");
    if (pt.n_l == 0 || pt.n_l > 0xffffffffULL) return set_err(ctx, XPG_EINVAL, "fused chain length outside 32 bits");
    if (prm->m > prm0->m) return set_err(ctx, XPG_EINVAL, "fused parts: the first part has the most moduli");
    for (uint32_t c = 0; c < prm->m; c++)
      if (prm->moduli[c] != prm0->moduli[c]) return set_err(ctx, XPG_EINVAL, "fused parts with other moduli");
    const uint32_t q4 = pt.qn >= (uint32_t)kFusedQR ? pt.qn - pt.qn % kFusedQR : 0;
    for (int half = 0; half < 2; half++) {
      const uint32_t qfirst = half ? q4 : 0, qn = half ? pt.qn - q4 : q4;
      if (qn == 0) continue;
      SegH h{};
      h.part = &pt;
      h.q_first = qfirst;
      h.s.db = pt.db;
      h.s.n_l = pt.n_l;
      h.s.groups = pt.groups;
      h.s.P = pt.P;
      h.s.n = prm->n;
      h.s.m = prm->m;
      h.s.sblocks = prm->n / 64;
      h.s.qn = qn;
      h.s.qr = half ? qn : (uint32_t)kFusedQR;
      h.s.qblocks = (qn + h.s.qr - 1) / h.s.qr;
      segs.push_back(h);
    }
    queries += pt.qn;
    steps += (uint64_t)pt.qn * pt.groups * pt.P * prm->m * prm->n * 2 * pt.n_l;
  }
  if (segs.size() > (size_t)kFusedSegs) return set_err(ctx, XPG_EINVAL, "fused launch of more than 24 segments");
  // longest tiles first in every XCD's list (a tile is W*GR rows x qr queries x n_l steps)
  std::stable_sort(segs.begin(), segs.end(),
                   [](const SegH &a, const SegH &b) { return a.s.n_l * a.s.qr > b.s.n_l * b.s.qr; });
  FusedArgs fa{};
  fa.ms = modset_of(prm0);
  fa.nseg = (uint32_t)segs.size();
  std::vector<uint64_t> image(8 + 2 * (size_t)queries, 0);  // 64 B of counters, then pointers
  uint64_t tiles[8] = {0};
  uint32_t ptr = 0;
  for (size_t i = 0; i < segs.size(); i++) {
    SegH &h = segs[i];
    const uint64_t colgroups = ((h.s.groups * h.s.P + W * GR - 1) / (W * GR)) * h.s.m * h.s.sblocks;
    for (int x = 0; x < 8; x++) {
      h.s.begin[x] = (uint32_t)tiles[x];
      tiles[x] += (colgroups > (uint64_t)x ? (colgroups - x + 7) / 8 : 0) * h.s.qblocks;
      if (tiles[x] > 0x7fffffffULL) return set_err(ctx, XPG_EINVAL, "fused grid too large");
    }
    h.s.ptr0 = ptr;
    for (uint32_t q = 0; q < h.s.qn; q++) {
      image[8 + ptr + q] = reinterpret_cast<uintptr_t>(h.part->q[h.q_first + q]);
      image[8 + ptr + h.s.qn + q] = reinterpret_cast<uintptr_t>(h.part->out[h.q_first + q]);
    }
    ptr += 2 * h.s.qn;
    fa.seg[i] = h.s;
  }
  uint64_t total = 0;
  for (int x = 0; x < 8; x++) {
    fa.tiles[x] = (uint32_t)tiles[x];
    total += tiles[x];
  }
  fa.counters = reinterpret_cast<uint32_t *>(dev_table);
  fa.ptrs = reinterpret_cast<const uint64_t *>(reinterpret_cast<char *>(dev_table) + 64);
  int rc = copy_h2d_table(ctx, dev_table, image.data(), image.size() * sizeof(uint64_t));
  if (rc) return rc;
  const uint32_t grid = std::max<uint32_t>(8, std::min<uint32_t>(fused_grid(ctx), (uint32_t)((total + 7) / 8 * 8)));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  timing_begin(ctx, &e0, &e1);
  hipEvent_t k0 = nullptr, k1 = nullptr;
  timing_begin(ctx, &k0, &k1, 2);
  if (ctx->timing && !ctx->ws) ctx->batch_steps += steps;
  hipLaunchKernelGGL((mac_fused_kernel<W>), dim3(grid), dim3(64 * W), 0, ctx->stream, fa);
  rc = hipGetLastError() != hipSuccess ? set_err(ctx, XPG_EDEVICE, "fused MAC launch") : XPG_OK;
  timing_end(ctx, k0, k1, 2);
  timing_end(ctx, e0, e1);
  return rc;
}

int launch_mac_level_multi(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *db, const uint64_t *const *q,
                           uint32_t nreq, uint64_t n_l, uint64_t gpr, uint32_t P, uint64_t *const *out_per_req,
                           uint64_t *out) {
  if (nreq == 0 || nreq > (uint32_t)kMaxReqs) return set_err(ctx, XPG_EINVAL, "1..64 requests per grouped launch");
  MultiArgs m{};
  m.base.ms = modset_of(prm);
  m.base.db = db;
  m.base.q = nullptr;
  m.base.out = out;
  m.base.n_l = n_l;
  m.base.groups = gpr * nreq;
  m.base.P = P;
  m.base.n = prm->n;
  m.base.m = prm->m;
  m.base.logn = prm->logn;
  for (uint32_t i = 0; i < nreq; i++) {
    m.q[i] = q[i];
    m.out[i] = out_per_req ? out_per_req[i] : nullptr;
  }
  m.gpr = gpr;
  m.per_req_out = out_per_req ? 1u : 0u;
  const uint64_t lanes = m.base.groups * P * (uint64_t)prm->m * prm->n;
  const uint64_t blocks = (lanes + 255) / 256;
  if (blocks > 0x7fffffffULL) return set_err(ctx, XPG_EINVAL, "MAC grid too large");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  timing_begin(ctx, &e0, &e1);
  hipEvent_t k0 = nullptr, k1 = nullptr;
  timing_begin(ctx, &k0, &k1, 1);
  // every poly of a group in one lane where the poly count has a build (a lane per group
  // and slot), else one chain pair per lane
  const uint64_t row_blocks = (m.base.groups * (uint64_t)prm->m * prm->n + 255) / 256;
  switch (P) {
    case 6: hipLaunchKernelGGL(mac_multi_rows_kernel<6>, dim3((uint32_t)row_blocks), dim3(256), 0, ctx->stream, m); break;
    case 12: hipLaunchKernelGGL(mac_multi_rows_kernel<12>, dim3((uint32_t)row_blocks), dim3(256), 0, ctx->stream, m); break;
    default: hipLaunchKernelGGL(mac_direct_multi_kernel, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, m); break;
  }
  timing_end(ctx, k0, k1, 1);
  timing_end(ctx, e0, e1);
  if (hipGetLastError() != hipSuccess) return set_err(ctx, XPG_EDEVICE, "grouped MAC launch");
  return XPG_OK;
}
