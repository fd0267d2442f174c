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

// One lane = both chains (a and b) of one (g, k, c, s): x is loaded once per element.
// Lanes with consecutive s read consecutive words, so every load is a coalesced 512 B
// row per wavefront.
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
  const uint64_t p = a.ms.p[c], delta = a.ms.delta[c];
  const uint64_t xstride = (uint64_t)a.P * mn, qstride = 2 * mn;
  const uint64_t *x = a.db + (g * a.n_l * a.P + k) * mn + cs;
  const uint64_t *ya = a.q + cs, *yb = a.q + mn + cs;
  uint64_t acc_a = 0, acc_b = 0;
  uint64_t e = 0;
  for (; e + 4 <= a.n_l; e += 4) {
    uint64_t xv[4], av[4], bv[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      xv[i] = __builtin_nontemporal_load(x + (e + i) * xstride);
      av[i] = ya[(e + i) * qstride];
      bv[i] = yb[(e + i) * qstride];
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      acc_a = mac_step(acc_a, xv[i], av[i], shoup_quot(av[i], p, delta), p);
      acc_b = mac_step(acc_b, xv[i], bv[i], shoup_quot(bv[i], p, delta), p);
    }
  }
  for (; e < a.n_l; e++) {
    const uint64_t xv = x[e * xstride], av = ya[e * qstride], bv = yb[e * qstride];
    acc_a = mac_step(acc_a, xv, av, shoup_quot(av, p, delta), p);
    acc_b = mac_step(acc_b, xv, bv, shoup_quot(bv, p, delta), p);
  }
  uint64_t *o = a.out + gk * 2 * mn + cs;
  o[0] = acc_a;
  o[mn] = acc_b;
}

}  // namespace

int launch_mac_level(xpg_ctx *ctx, const xpg_params *prm, const uint64_t *db, const uint64_t *q,
                     uint64_t n_l, uint64_t groups, uint32_t P, uint64_t *out) {
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
  const uint64_t lanes = groups * P * (uint64_t)prm->m * prm->n;
  const uint64_t blocks = (lanes + 255) / 256;
  if (blocks > 0x7fffffffULL) return set_err(ctx, XPG_EINVAL, "MAC grid too large");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  timing_begin(ctx, &e0, &e1);
  hipLaunchKernelGGL(mac_direct_kernel, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, a);
  if (hipGetLastError() != hipSuccess) return set_err(ctx, XPG_EDEVICE, "MAC launch");
  timing_end(ctx, e0, e1);
  return XPG_OK;
}
