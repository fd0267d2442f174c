// mx.hip -- level 0 of the reply on the matrix cores: the residues of the chains' sums by
// int8 MFMA over signed byte digits, the chains' last steps by the reference's own step.
//
// Reference path: the level-0 chains of PIRReplyGeneratorNFL_internal::generateReply
// (pir/replyGenerator/PIRReplyGeneratorNFL_internal.cpp:511-595) through
// NFLlib::mulandaddShoup (crypto/NFLlib.hpp:303-321).  A chain (query q, half s, unit u =
// group * P + poly, slot sl = modulus k * n + coefficient c) runs, from acc = 0, over the
// elements e = 0 .. n_l - 1:
//     acc = acc + r_e,  r_e = x_e*y_e - hi64(x_e*y'_e)*p  (in [0, 2p));  if (acc >= p) acc -= p.
// Write acc = rho + j*p (rho < p) and r_e = u_e + c_e*p (u_e = x_e*y_e mod p).  Then rho is
// the residue of the sum of x*y over the elements so far, and j evolves by
// j' = max(j + b_e + c_e - 1, 0) with b_e the residue wrap [rho + u_e >= p] (the map the
// streaming MAC in mac.hip composes).  j stays at most 8 (acc < 2^64 < 9p).  So:
//   1. The residue of the first E0 = n_l - 40 elements is an exact integer dot product
//      reduced mod p: the matrix cores compute it (over whole blocks of 32 elements, the
//      query digits past E0 zero).
//   2. The tail's steps run with the reference's step from acc = that residue (j = 0),
//      counting the steps that do not subtract p (`kept`).  Over the tail a start at j0
//      ends at max(j0 + D, M) and the start at 0 at M, with D = M - kept, so every start
//      j0 <= kept ends alike: kept >= 8 covers every true start (mac.hip: tail_word).
//   3. Chains whose tail does not settle this (bit 63 of their word), and every chain of a
//      query with a word not below its modulus (outside the model: its r_e may leave
//      [0, 2p)), are recomputed whole with the reference's step (mx_fixup_kernel).
// The words are the reference's: same residue, same excess multiple of p.
//
// Digits: a word x < 2^63 - 2^56 is x = sum_j d_j 256^j with d_j in [-128, 127] the bytes
// of (x + B) ^ B, B = 0x8080808080808080 (the bytes of x + B are d_j + 128).  For a slot,
//     sum_e y_e x_e = sum_{i,j} 2^(8(i+j)) S_ij,   S_ij = sum_e dy_i(e) dx_j(e),
// and v_mfma_i32_32x32x32_i8 accumulates S_ij for row (query pair, half, digit i) and
// column (unit of four, digit j), 32 elements per instruction, exactly in int32 (|dy dx| <=
// 2^14).  Prefixes are cut into passes of 2^15 elements, so |S_ij| <= 2^29 and the
// epilogue may pair two lanes' sums exactly in 64 bits before reducing mod p.
//
// Fragments (1 KiB, one instruction operand): lane L holds row (or column) L % 32 and the
// 16 elements 16*(L/32) .. +15 of a block of 32, one byte each (the instruction pairs A's
// and B's bytes by lane half and byte position).  The database's fragments, laid out once
// when it is imported ("matrix layout"), are [slot][unit tile of 4][element block]; a
// launch lays its queries' out as [slot][row tile of 2 queries][element block].
#include <algorithm>

#include "internal.hpp"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr uint64_t kDigitBias = 0x8080808080808080ull;
constexpr uint64_t kDigitMax = 0x7F7F7F7F7F7F7F80ull;  // words below it have 8 signed digits
constexpr uint64_t kM61 = (1ull << 61) - 1;
constexpr uint32_t kTailMin = 40;                      // fewest tail steps
constexpr uint32_t kPassBlocks = 1024;                 // element blocks per pass (2^15 elements)
constexpr int kNT = 2;                                 // column tiles per wave
constexpr int kStageBlocks = 2;                        // element blocks per LDS stage

__device__ __forceinline__ uint64_t digits_of(uint64_t x) { return (x + kDigitBias) ^ kDigitBias; }

// 4x4 byte transpose: y_k = (x0.b_k, x1.b_k, x2.b_k, x3.b_k)
__device__ __forceinline__ void bt4(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3, uint32_t (&y)[4]) {
  const uint32_t a = __builtin_amdgcn_perm(x1, x0, 0x05010400u), b = __builtin_amdgcn_perm(x1, x0, 0x07030602u);
  const uint32_t c = __builtin_amdgcn_perm(x3, x2, 0x05010400u), d = __builtin_amdgcn_perm(x3, x2, 0x07030602u);
  y[0] = __builtin_amdgcn_perm(c, a, 0x05040100u);
  y[1] = __builtin_amdgcn_perm(c, a, 0x07060302u);
  y[2] = __builtin_amdgcn_perm(d, b, 0x05040100u);
  y[3] = __builtin_amdgcn_perm(d, b, 0x07060302u);
}

// 16 digit words w[t] -> the eight 16-byte lane chunks of one (row or unit, lane half):
// chunk j holds digit j of w[0..15], written to dst + 16*j.
__device__ __forceinline__ void store_chunks(const uint64_t (&w)[16], uint8_t *dst) {
  uint32_t o[8][4];  // [digit][dword]
#pragma unroll
  for (int q = 0; q < 4; q++) {
#pragma unroll
    for (int half = 0; half < 2; half++) {
      uint32_t y[4];
      bt4((uint32_t)(w[4 * q] >> (32 * half)), (uint32_t)(w[4 * q + 1] >> (32 * half)),
          (uint32_t)(w[4 * q + 2] >> (32 * half)), (uint32_t)(w[4 * q + 3] >> (32 * half)), y);
#pragma unroll
      for (int k = 0; k < 4; k++) o[4 * half + k][q] = y[k];
    }
  }
#pragma unroll
  for (int j = 0; j < 8; j++) {
    v4i v = {(int)o[j][0], (int)o[j][1], (int)o[j][2], (int)o[j][3]};
    *reinterpret_cast<v4i *>(dst + 16 * j) = v;
  }
}

// ---- the database in the matrix layout ----
struct PackDbArgs {
  const uint64_t *words;  // row-major: ((g*n_l + e)*P + pl)*m*n + sl
  uint8_t *frag;
  uint32_t *bad;          // set when a word has no 8-digit form
  uint64_t n_l, total;    // threads
  uint32_t S, U, UT, EB, P;
};

// thread (sl fastest, lane half h, unit u of the padded tiles, element block eb): digits of
// its 16 words into lanes (u % 4) * 8 + j + 32 h of fragment (sl, u / 4, eb)
__global__ void __launch_bounds__(256) mx_pack_db_kernel(PackDbArgs a) {
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= a.total) return;
  const uint32_t sl = (uint32_t)(t % a.S);
  uint64_t r = t / a.S;
  const uint32_t h = (uint32_t)(r & 1);
  r >>= 1;
  const uint32_t u = (uint32_t)(r % (a.UT * 4));
  const uint32_t eb = (uint32_t)(r / (a.UT * 4));
  const uint64_t mn = a.S;
  uint64_t w[16];
  bool bad = false;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const uint64_t e = (uint64_t)eb * 32 + 16 * h + k;
    uint64_t x = 0;
    if (u < a.U && e < a.n_l) {
      const uint64_t g = u / a.P, pl = u % a.P;
      x = __builtin_nontemporal_load(a.words + ((g * a.n_l + e) * a.P + pl) * mn + sl);
    }
    bad |= x >= kDigitMax;
    w[k] = digits_of(x);
  }
  if (bad) *a.bad = 1;
  uint8_t *dst = a.frag + (((uint64_t)sl * a.UT + u / 4) * a.EB + eb) * 1024 + 16 * ((u % 4) * 8 + 32 * h);
  store_chunks(w, dst);
}

// dst[g*per + i] = src[g*stride + i] for i < per (each group's last elements)
__global__ void __launch_bounds__(256) mx_tail_copy_kernel(const uint64_t *src, uint64_t *dst, uint64_t per,
                                                          uint64_t stride, uint64_t total) {
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= total) return;
  dst[t] = src[(t / per) * stride + t % per];
}

// ---- the queries of one launch ----
constexpr int kMaxQ = 64;

struct PackQArgs {
  const uint64_t *q[kMaxQ];  // level-0 query words: (e*2 + s)*m*n + sl
  uint8_t *frag;
  uint32_t *qbad;            // [kMaxQ]: a word of the query is not below its modulus
  ModSet ms;
  uint64_t n_l, E0, total;
  uint32_t S, n, qn, RTp, EB, EBP;
};

// thread (sl fastest, h, half s, query row qq < 2*RTp, element block eb < EB): digits of its
// 16 words into lanes (qq % 2) * 16 + s * 8 + j + 32 h of fragment (sl, qq / 2, eb) for the
// prefix blocks eb < EBP (elements from E0 on as zero digits); every element's word is
// checked against the modulus
__global__ void __launch_bounds__(256) mx_pack_q_kernel(PackQArgs a) {
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= a.total) return;
  const uint32_t sl = (uint32_t)(t % a.S);
  uint64_t r = t / a.S;
  const uint32_t h = (uint32_t)(r & 1), s = (uint32_t)((r >> 1) & 1);
  r >>= 2;
  const uint32_t qq = (uint32_t)(r % (a.RTp * 2));
  const uint32_t eb = (uint32_t)(r / (a.RTp * 2));
  const uint64_t p = a.ms.p[sl / a.n];
  const uint64_t mn = a.S;
  uint64_t w[16];
  bool bad = false;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const uint64_t e = (uint64_t)eb * 32 + 16 * h + k;
    uint64_t y = 0;
    if (qq < a.qn && e < a.n_l) y = __builtin_nontemporal_load(a.q[qq] + (e * 2 + s) * mn + sl);
    bad |= y >= p;
    w[k] = e < a.E0 ? digits_of(y) : 0;
  }
  if (bad) a.qbad[qq] = 1;
  if (eb >= a.EBP) return;
  uint8_t *dst = a.frag + (((uint64_t)sl * a.RTp + qq / 2) * a.EBP + eb) * 1024 + 16 * ((qq % 2) * 16 + s * 8 + 32 * h);
  store_chunks(w, dst);
}

__global__ void mx_clear_kernel(uint32_t *p, uint32_t n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0;
}

// ---- the matrix-core MAC ----
//
// A workgroup of four waves takes an item (slot, column block cb, row block rb): MT row
// tiles shared by the four waves and 4 x NT column tiles, NT per wave (tiles of 32 x 32).
// Per stage of two blocks of 32 elements the row fragments go through LDS (wave w loads
// row fragment w) and each wave loads its own column fragments straight into registers:
// LDS carries only what several waves read (its bandwidth, not the matrix cores, bounded
// a layout whose waves shared both kinds).  Items run rb fastest: the row blocks of one
// column block follow one another on one XCD (its database fragments come from HBM
// once), and an XCD's items stay within few slots (their query fragments stay in its L2).
struct MacArgsMx {
  const uint8_t *A;  // query fragments [S][RTp][EBP]
  const uint8_t *B;  // database fragments [S][UT][EB]
  uint64_t *stage;   // per item 4 * MT*NT*16 words: (wave, tile i, tile j, row group, unit)
  ModSet ms;
  uint64_t items, per_xcd;
  uint32_t n, RTp, UT, EB, EBP, RB, CB;
};

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
  const int lo = __shfl_xor((int)(uint32_t)v, m, 64), hi = __shfl_xor((int)(uint32_t)(v >> 32), m, 64);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// lane exchanges within quads and half rows by DPP (one VALU instruction per dword)
constexpr int kQuadXor1 = 0xB1, kQuadXor2 = 0x4E, kQuadRev = 0x1B, kHalfMirror = 0x141;

template <int CTRL>
__device__ __forceinline__ uint64_t dpp64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)v, CTRL, 0xF, 0xF, true);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, true);
  return ((uint64_t)hi << 32) | lo;
}

// a * b + c (signed 32 x 32 + 64 bits), one v_mad_i64_i32
__device__ __forceinline__ uint64_t mad_i64(int32_t a, int32_t b, uint64_t c) {
  uint64_t d, carry;  // the carry-out goes to a scalar pair of its own (never read)
  asm("v_mad_i64_i32 %0, %1, %2, %3, %4" : "=v"(d), "=s"(carry) : "v"(a), "s"(b), "v"(c));
  return d;
}

// v mod p for v < 2p, with negp = 2^64 - p in scalar registers (an add, a compare, two selects)
__device__ __forceinline__ uint64_t canon(uint64_t v, uint64_t p, uint64_t negp) {
  const uint64_t t = v + negp;
  return v >= p ? t : v;
}

// a wave-uniform 64-bit value in scalar registers
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Epilogue of one tile: lane L holds S for column (unit ul = (L>>3)&3, digit lx = L&7) and
// rows 8g + 4h + b (g: query of the pair and half; digits 4h + b), h = L>>5, in acc[4g+b].
// Returns, in the lanes with h = 0 and lx < 4, the output of row group g = lx for unit ul:
// sum over the 16 lanes of unit ul of 2^(8(4h+lx)) * sum_b 2^(8b) S, reduced mod p.
//   D_g = sum_b 2^(8b) S (|D| < 2^53.01: |S| <= 2^29 for passes of 2^15 elements);
//   xor 1, exact: the even lane keeps groups 0 and 2, the odd lane 1 and 3, each as
//     P = D(even lane) + 2^8 D(odd lane)  (|P| < 2^61.02);
//   weights: P + 4p (positive) times 2^(8(4h + (lx & ~1))) mod p, a Shoup product;
//   xor 2: the lanes with bit 1 clear keep groups {0, 1}, the others {2, 3} (so lane lx
//     holds group lx & 3); xor 4 and xor 32 add the partners' partial sums mod p.
// xor 1, 2 and 4 are DPP moves (xor 4 = quad reversal, then half-row mirror); xor 32 a
// permute through LDS.
__device__ __forceinline__ uint64_t tile_epilogue(const v16i &acc, uint32_t lane, uint64_t p, uint64_t negp,
                                                  uint32_t delta, uint64_t pw, uint64_t pwq, uint64_t p4) {
  uint64_t D[4];
#pragma unroll
  for (int g = 0; g < 4; g++)
    D[g] = mad_i64(acc[4 * g + 3], 1 << 24,
                   mad_i64(acc[4 * g + 2], 1 << 16, mad_i64(acc[4 * g + 1], 1 << 8, mad_i64(acc[4 * g], 1, 0))));
  const bool odd = lane & 1, hi2 = (lane >> 1) & 1;
  const uint32_t sk = odd ? 8 : 0, sr = 8 - sk;  // P = (even lane's D) + 2^8 (odd lane's D)
  // xor 1: send the groups the partner keeps
  const uint64_t r0 = dpp64<kQuadXor1>(odd ? D[0] : D[1]);
  const uint64_t r1 = dpp64<kQuadXor1>(odd ? D[2] : D[3]);
  const uint64_t P0 = ((odd ? D[1] : D[0]) << sk) + (r0 << sr);  // group odd
  const uint64_t P1 = ((odd ? D[3] : D[2]) << sk) + (r1 << sr);  // group 2 + odd
  // weighted, in [0, 2p)
  const uint64_t w0 = shoup_lazy_asm(P0 + p4, pw, pwq, delta);
  const uint64_t w1 = shoup_lazy_asm(P1 + p4, pw, pwq, delta);
  // xor 2 (keep group odd when bit 1 is clear, else 2 + odd) and xor 4: below 8p < 2^64
  uint64_t v = (hi2 ? w1 : w0) + dpp64<kQuadXor2>(hi2 ? w0 : w1);
  v += dpp64<kHalfMirror>(dpp64<kQuadRev>(v));
  v = canon((v & kM61) + (v >> 61) * delta, p, negp);
  return canon(v + shfl_xor64(v, 32), p, negp);
}

template <int MT>
__global__ void __launch_bounds__(256, 2) mx_mac_kernel(MacArgsMx a) {
  constexpr int NT = kNT, TR = MT, TC = 4 * NT, KB = kStageBlocks;
  __shared__ v4i lds[2][KB][MT][64];
  const uint32_t bid = blockIdx.x;
  const uint64_t item = (uint64_t)(bid & 7) * a.per_xcd + (bid >> 3);
  if (item >= a.items) return;  // the whole workgroup
  const uint32_t rb = (uint32_t)(item % a.RB);
  const uint64_t t = item / a.RB;
  const uint32_t cb = (uint32_t)(t % a.CB);
  const uint32_t sl = (uint32_t)(t / a.CB);
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t k = sl / a.n;
  const uint64_t p = a.ms.p[k];
  const uint32_t delta = a.ms.delta[k];
  // wave w stages row fragment w (when w < MT) and loads its own NT column fragments
  // (column tiles past the database read its last tile, whose results are dropped), each
  // from a wave-uniform base plus the lane's 32-bit offset
  const bool stager = MT == 4 || wave < (uint32_t)MT;
  const uint8_t *srcA = a.A + uniform64(((uint64_t)sl * a.RTp + rb * TR + min(wave, (uint32_t)MT - 1)) * a.EBP * 1024);
  const uint8_t *srcB[NT];
#pragma unroll
  for (int j = 0; j < NT; j++)
    srcB[j] = a.B + uniform64(((uint64_t)sl * a.UT + min(cb * TC + wave * NT + j, a.UT - 1)) * a.EB * 1024);
  // a stage is KB consecutive blocks (indices past `last` read block `last` again)
  auto loadA = [&](v4i (&r)[KB], uint32_t eb, uint32_t last) {
    if (stager) {
#pragma unroll
      for (int kb = 0; kb < KB; kb++)
        r[kb] = *reinterpret_cast<const v4i *>(srcA + min(eb + kb, last) * 1024 + lane * 16);
    }
  };
  auto putA = [&](int buf, const v4i (&r)[KB]) {
    if (stager) {
#pragma unroll
      for (int kb = 0; kb < KB; kb++) lds[buf][kb][wave][lane] = r[kb];
    }
  };
  auto loadB = [&](v4i (&r)[KB][NT], uint32_t eb, uint32_t last) {
#pragma unroll
    for (int kb = 0; kb < KB; kb++) {
      const uint32_t off = min(eb + kb, last) * 1024 + lane * 16;
#pragma unroll
      for (int j = 0; j < NT; j++) r[kb][j] = *reinterpret_cast<const v4i *>(srcB[j] + off);
    }
  };
  // the epilogue's per-lane weight 2^(8(4h + (lx & ~1))) mod p and its Shoup quotient
  const uint32_t wexp = 8 * (4 * (lane >> 5) + (lane & 6));
  const uint64_t pw = wexp < 64 ? (1ull << wexp) : ((uint64_t)delta << (wexp - 61));
  const uint64_t pwq = shoup_quot(pw, p, delta);
  uint64_t *st = a.stage + (item * 4 + wave) * (MT * NT * 16);
  for (uint32_t c0 = 0; c0 < a.EBP; c0 += kPassBlocks) {
    const uint32_t c1 = min(a.EBP, c0 + kPassBlocks);
    v16i acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
      for (int j = 0; j < NT; j++) acc[i][j] = v16i{};
    // Row fragments: two register sets run two stages ahead of the LDS buffer being read
    // (at the top of stage eb, LDS buffer ((eb - c0) / KB) & 1 holds it, one set the next
    // stage and the other the one after).  Column fragments: two register sets, the
    // current stage's and the next's, each reloaded two stages ahead once its products
    // are issued.  Block indices past the pass are clamped to its last block (loads
    // without branches, so each wait names only the older loads); such copies are never
    // multiplied.
    const uint32_t last = c1 - 1;
    v4i RA0[KB], RA1[KB], B0[KB][NT], B1[KB][NT];
    loadA(RA0, c0, last);
    loadB(B0, c0, last);
    putA(0, RA0);
    loadA(RA0, c0 + KB, last);
    loadA(RA1, c0 + 2 * KB, last);
    loadB(B1, c0 + KB, last);
    __syncthreads();
    for (uint32_t eb = c0; eb < c1; eb += 2 * KB) {
#pragma unroll
      for (int half = 0; half < 2; half++) {
        const uint32_t e0 = eb + half * KB;
        v4i (&B)[KB][NT] = half ? B1 : B0;
#pragma unroll
        for (int kb = 0; kb < KB; kb++) {
          if (e0 + kb < c1) {
            v4i fa[MT];
#pragma unroll
            for (int i = 0; i < MT; i++) fa[i] = lds[half][kb][i][lane];
#pragma unroll
            for (int i = 0; i < MT; i++)
#pragma unroll
              for (int j = 0; j < NT; j++)
                acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[i], B[kb][j], acc[i][j], 0, 0, 0);
          }
        }
        putA(half ^ 1, half ? RA1 : RA0);
        loadA(half ? RA1 : RA0, e0 + 3 * KB, last);
        loadB(B, e0 + 2 * KB, last);
        __syncthreads();
      }
    }
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
      for (int j = 0; j < NT; j++) {
        uint64_t v = tile_epilogue(acc[i][j], lane, p, 0 - p, delta, pw, pwq, 4 * p);
        if (lane < 32 && (lane & 7) < 4) {
          uint64_t *o = st + ((i * NT + j) * 4 + (lane & 7)) * 4 + ((lane >> 3) & 3);
          if (c0 > 0) v = canon(v + *o, p, 0 - p);
          *o = v;
        }
      }
  }
}

// ---- whole chains where the tail did not settle them ----
struct FixArgs {
  const uint8_t *frag;       // database fragments [S][UT][EB]
  const uint64_t *q[kMaxQ];  // level-0 query words
  uint64_t *out[kMaxQ];
  const uint32_t *qbad;
  ModSet ms;
  uint64_t n_l, total;
  uint32_t S, n, U, UT, EB, P;
};

// block: 1024 consecutive slots of one output row (query q, unit u, half s), four per
// thread; the row index is decoded once per block.  A marked word, or
// every word of a query with a word not below its modulus, takes the whole chain.
constexpr uint32_t kFixSlots = 1024;

__global__ void __launch_bounds__(256) mx_fixup_kernel(FixArgs a) {
  const uint32_t blocks_per_row = (a.S + kFixSlots - 1) / kFixSlots;
  const uint64_t row = blockIdx.x / blocks_per_row;
  const uint32_t sl0 = (uint32_t)(blockIdx.x % blocks_per_row) * kFixSlots + threadIdx.x * 4;
  if (sl0 >= a.S) return;
  const uint32_t s = (uint32_t)(row & 1);
  const uint32_t u = (uint32_t)((row >> 1) % a.U), q = (uint32_t)((row >> 1) / a.U);
  const uint64_t mn = a.S;
  uint64_t *o = a.out[q] + ((uint64_t)u * 2 + s) * mn + sl0;
  const bool all = a.qbad[q] != 0;
  uint64_t w[4];  // (caller buffers are 8-byte aligned: four word loads)
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] = sl0 + i < a.S ? o[i] : 0;
  if (!((w[0] | w[1] | w[2] | w[3]) >> 63) && !all) return;
  for (int i = 0; i < 4; i++) {
    if (sl0 + i >= a.S || (!(w[i] >> 63) && !all)) continue;
    const uint32_t sl = sl0 + i;
    const uint32_t k = sl / a.n;
    const uint64_t p = a.ms.p[k];
    const uint32_t delta = a.ms.delta[k];
    const uint8_t *f0 = a.frag + ((uint64_t)sl * a.UT + u / 4) * a.EB * 1024 + 16 * ((u % 4) * 8);
    const uint64_t *y = a.q[q] + (uint64_t)s * mn + sl;
    uint64_t acc = 0;
    for (uint64_t e = 0; e < a.n_l; e++) {
      const uint8_t *f = f0 + (e / 32) * 1024 + 16 * 32 * ((e % 32) / 16) + (e % 16);
      uint64_t b = 0;
#pragma unroll
      for (int j = 0; j < 8; j++) b |= (uint64_t)f[16 * j] << (8 * j);
      const uint64_t x = (b ^ kDigitBias) - kDigitBias;
      const uint64_t yv = y[e * 2 * mn];
      acc = mac_step(acc, x, yv, shoup_quot(yv, p, delta), p, delta);
    }
    o[i] = acc;
  }
}

uint64_t al256(uint64_t b) { return (b + 255) & ~(uint64_t)255; }

// workgroup shape by row tiles (query pairs): WR waves along the rows, MT row tiles each
struct WgShape {
  int WR, MT;
  uint32_t TR, TC;
};
WgShape shape_for(uint32_t qn) {
  const uint32_t RT = (qn + 1) / 2;
  const int MT = RT > 2 ? 4 : (int)RT;
  return WgShape{1, MT, (uint32_t)MT, (uint32_t)(4 * kNT)};
}

}  // namespace

bool mx_geometry(const xpg_params *prm, uint64_t n_l, uint64_t groups, uint32_t P, MxGeom *g) {
  // slots in whole 256-lane blocks; 32-bit fragment offsets
  if (prm->n % 64 != 0 || (prm->m * prm->n) % 256 != 0 || n_l < kTailMin + 32 || n_l >= (1ull << 27)) return false;
  g->S = prm->m * prm->n;
  g->U = (uint32_t)(groups * P);
  g->UT = (g->U + 3) / 4;
  g->EB = (uint32_t)((n_l + 31) / 32);
  g->E0 = n_l - kTailMin;
  g->EBP = (uint32_t)((g->E0 + 31) / 32);
  g->L = kTailMin;
  g->n_l = n_l;
  g->groups = groups;
  g->P = P;
  return groups * P < (1ull << 31);
}

uint64_t mx_frag_bytes(const MxGeom &g) { return (uint64_t)g.S * g.UT * g.EB * 1024; }

uint64_t mx_tail_words(const MxGeom &g) { return g.groups * g.L * g.P * (uint64_t)g.S; }

int launch_mx_pack_db(xpg_ctx *ctx, const xpg_params *prm, const MxGeom &g, const uint64_t *words, uint8_t *frag,
                      uint64_t *tail, uint32_t *bad_dev) {
  PackDbArgs a{};
  a.words = words;
  a.frag = frag;
  a.bad = bad_dev;
  a.n_l = g.n_l;
  a.S = g.S;
  a.U = g.U;
  a.UT = g.UT;
  a.EB = g.EB;
  a.P = g.P;
  a.total = (uint64_t)g.S * g.UT * 4 * 2 * g.EB;
  const uint64_t blocks = (a.total + 255) / 256;
  if (blocks > 0x7fffffffULL) return set_err(ctx, XPG_EINVAL, "database too large for the matrix layout");
  hipLaunchKernelGGL(mx_pack_db_kernel, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, a);
  if (hipGetLastError() != hipSuccess) return set_err(ctx, XPG_EDEVICE, "matrix layout launch");
  // each group's last L elements, row-major, for the chains' tails
  const uint64_t row = (uint64_t)g.P * g.S, per_group = g.L * row, tw = g.groups * per_group;
  const uint64_t tblocks = (tw + 255) / 256;
  if (tblocks > 0x7fffffffULL) return set_err(ctx, XPG_EINVAL, "database too large for the matrix layout");
  hipLaunchKernelGGL(mx_tail_copy_kernel, dim3((uint32_t)tblocks), dim3(256), 0, ctx->stream, words + g.E0 * row,
                     tail, per_group, g.n_l * row, tw);
  if (hipGetLastError() != hipSuccess) return set_err(ctx, XPG_EDEVICE, "tail rows copy");
  (void)prm;
  return XPG_OK;
}

// workspace of one launch of qn queries: query fragments, staged residues, query flags
static void mx_work_parts(const MxGeom &g, uint32_t qn, uint64_t *a_bytes, uint64_t *st_bytes) {
  const WgShape w = shape_for(qn);
  const uint32_t RT = (qn + 1) / 2, RB = (RT + w.TR - 1) / w.TR, RTp = RB * w.TR;
  const uint32_t CB = (g.UT + w.TC - 1) / w.TC;
  *a_bytes = (uint64_t)g.S * RTp * g.EBP * 1024;
  *st_bytes = (uint64_t)g.S * CB * RB * 4 * w.MT * kNT * 16 * 8;
}

uint64_t mx_work_bytes(const MxGeom &g, uint32_t qn) {
  uint64_t ab = 0, sb = 0;
  mx_work_parts(g, qn, &ab, &sb);
  return al256(ab) + al256(sb) + 256;
}

int launch_mx_level(xpg_ctx *ctx, const xpg_params *prm, const MxGeom &g, const uint8_t *frag, const uint64_t *tail,
                    const uint64_t *const *q, uint64_t *const *out, uint32_t qn, void *work) {
  if (qn == 0 || qn > (uint32_t)kMaxQ) return set_err(ctx, XPG_EINVAL, "batch of 1..64 queries per launch");
  const WgShape w = shape_for(qn);
  const uint32_t RT = (qn + 1) / 2, RB = (RT + w.TR - 1) / w.TR, RTp = RB * w.TR;
  const uint32_t CB = (g.UT + w.TC - 1) / w.TC;
  uint64_t ab = 0, sb = 0;
  mx_work_parts(g, qn, &ab, &sb);
  uint8_t *A = static_cast<uint8_t *>(work);
  uint64_t *stage = reinterpret_cast<uint64_t *>(A + al256(ab));
  uint32_t *qbad = reinterpret_cast<uint32_t *>(A + al256(ab) + al256(sb));
  const ModSet ms = modset_of(prm);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  timing_begin(ctx, &e0, &e1);
  hipLaunchKernelGGL(mx_clear_kernel, dim3(1), dim3(64), 0, ctx->stream, qbad, (uint32_t)kMaxQ);
  {
    PackQArgs a{};
    for (uint32_t i = 0; i < qn; i++) a.q[i] = q[i];
    a.frag = A;
    a.qbad = qbad;
    a.ms = ms;
    a.n_l = g.n_l;
    a.E0 = g.E0;
    a.S = g.S;
    a.n = prm->n;
    a.qn = qn;
    a.RTp = RTp;
    a.EB = g.EB;
    a.EBP = g.EBP;
    a.total = (uint64_t)g.S * 4 * RTp * 2 * g.EB;
    const uint64_t blocks = (a.total + 255) / 256;
    if (blocks > 0x7fffffffULL) return set_err(ctx, XPG_EINVAL, "query layout grid too large");
    hipLaunchKernelGGL(mx_pack_q_kernel, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, a);
  }
  {
    MacArgsMx a{};
    a.A = A;
    a.B = frag;
    a.stage = stage;
    a.ms = ms;
    a.n = prm->n;
    a.RTp = RTp;
    a.UT = g.UT;
    a.EB = g.EB;
    a.EBP = g.EBP;
    a.RB = RB;
    a.CB = CB;
    a.items = (uint64_t)g.S * CB * RB;
    a.per_xcd = (a.items + 7) / 8;
    const uint64_t blocks = a.per_xcd * 8;
    if (blocks > 0x7fffffffULL) return set_err(ctx, XPG_EINVAL, "matrix MAC grid too large");
    hipEvent_t k0 = nullptr, k1 = nullptr;
    timing_begin(ctx, &k0, &k1, 2);
    // the timed kernel's chain steps: the prefixes (the tails run after it)
    if (ctx->timing && !ctx->ws) ctx->batch_steps += (uint64_t)qn * g.U * g.S * 2 * g.E0;
    const dim3 grid((uint32_t)blocks), wg(256);
    if (w.MT == 4) hipLaunchKernelGGL((mx_mac_kernel<4>), grid, wg, 0, ctx->stream, a);
    else if (w.MT == 2) hipLaunchKernelGGL((mx_mac_kernel<2>), grid, wg, 0, ctx->stream, a);
    else hipLaunchKernelGGL((mx_mac_kernel<1>), grid, wg, 0, ctx->stream, a);
    timing_end(ctx, k0, k1, 2);
  }
  if (hipGetLastError() != hipSuccess) return set_err(ctx, XPG_EDEVICE, "matrix MAC launch");
  {
    std::vector<const uint64_t *> qt(qn);
    for (uint32_t i = 0; i < qn; i++) qt[i] = q[i] + g.E0 * 2 * (uint64_t)g.S;
    // the tails start from the staged residues (no pass of their own to the outputs)
    TailStage ts;
    ts.stage = stage;
    ts.cb = CB;
    ts.rb = RB;
    ts.mt = (uint32_t)w.MT;
    const int rc = launch_mac_tail(ctx, prm, tail, qt.data(), out, qn, g.L, g.groups, g.P, ts, 0);
    if (rc) return rc;
  }
  {
    FixArgs a{};
    a.frag = frag;
    for (uint32_t i = 0; i < qn; i++) {
      a.q[i] = q[i];
      a.out[i] = out[i];
    }
    a.qbad = qbad;
    a.ms = ms;
    a.n_l = g.n_l;
    a.S = g.S;
    a.n = prm->n;
    a.U = g.U;
    a.UT = g.UT;
    a.EB = g.EB;
    a.P = g.P;
    a.total = (uint64_t)qn * g.U * 2 * g.S;
    const uint64_t blocks = (uint64_t)qn * g.U * 2 * ((g.S + kFixSlots - 1) / kFixSlots);
    if (blocks > 0x7fffffffULL) return set_err(ctx, XPG_EINVAL, "fixup grid too large");
    hipLaunchKernelGGL(mx_fixup_kernel, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, a);
  }
  const int rc = hipGetLastError() != hipSuccess ? set_err(ctx, XPG_EDEVICE, "matrix MAC launch") : XPG_OK;
  timing_end(ctx, e0, e1);
  return rc;
}
