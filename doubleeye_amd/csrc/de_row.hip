// The row pass of the dense path (stage 2: graded score + H-fwd + H-bwd of
// one image row) on gfx950, in two forms: the register-resident pass k_row
// and the streaming pass k_rows, with their launchers.
//
// Reference semantics (paths relative to the reference repository root):
//   graded score    core/tools/de_dense.cpp:1201-1243
//   int16 filter    core/tools/de_dense.cpp:411-430, 485-508 (H-fwd, H-bwd)
// Every value is bit-identical to the reference (de_common.h gives the
// offset form of the filter step).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <utility>

#include "de_batch.h"
#include "de_common.h"
#include "de_kern.h"

namespace de {

// ---- the row pass (k_row) ----
//
// One workgroup per image row, one wave per 64-disparity group, each wave one
// chain per lane (lane = disparity). The row's left records, right records
// (census and kTblRow * R) and the score table are staged in LDS once for the
// workgroup. Lanes read the left record of column x (a broadcast), the right
// record of column x - d (one 12-byte read) and the score table.
//
// The forward scan must be complete before the backward scan starts, so every
// forward value of the row is kept until the backward scan reaches it, in the
// home of its 8-column block: the last kRowVBlocks aligned blocks in VGPRs,
// the kRowABlocks before them in AGPRs (v_accvgpr_write / read: the
// accumulation registers are otherwise unused here), earlier blocks in the
// wave's LDS slots (as many as the launcher's LDS budget allows, the
// columns past the aligned range in the first kRowTailSlots of them) and the
// rest in a global scratch private to the (row, group) wave, [block][lane][8]
// u16. A wave then holds about 850 columns on chip: waves run one per SIMD
// (up to 512 registers each), one round for an 848x480x128 frame.
//
// Forward step (offset form, as in de_common.h) with the record's filter word
// wz = 2a << 16 | min(65536 - 2a, 65535):
//     acc = lo16(wz) * c' + 32768                 v_mad_u32_u16
//     t2  = hi16(wz) * hi16(t2) + acc             v_mad_u32_u16, op_sel
// For a = 0 the capped weight gives 65535 c' + 32768, whose high half is c'
// as 65536 c' + 32768's is (c' <= 32768), and only the high half is carried.
// The backward step reads its forward value from the high or low half of a
// packed word with op_sel: two multiply-adds a column.
//
// The forward scan is software-pipelined over 8-column blocks in three
// stages (LDS loads of block n+2, score-table reads of block n+1, chain of
// block n), from the block of the wave's first valid column to the row's end:
// the ragged band (lane d's window opens at column hw + d) and the columns
// past the aligned range take masked scores (kOff, a zero score, outside the
// window), so no column runs alone. The backward scan's final values leave
// as one 2-byte store a lane a column (global_store_short_d16_hi of the chain
// word), a wave's store covering the 128 contiguous bytes of its disparity
// group in the [y][x][Dpad] volume: no LDS round trip stands between the chain
// and the store (an LDS transpose into 16-byte stores took 71.7 us against
// 69.2 us at 848x480x128).

constexpr int kRowVBlocks = 16;      // forward blocks in VGPRs (64 registers)
constexpr int kRowABlocks = 60;      // forward blocks in AGPRs (240 registers)
constexpr int kRowRegBlocks = kRowVBlocks + kRowABlocks;
constexpr int kRowTailSlots = 3;     // LDS slots of the blocks from the end of the aligned range
constexpr int kRowPad = 16;          // LDS record arrays run this far past W

// Graded score of one column: census distance h (two popcounts), intensity
// offset kTblRow * min(|L - R|, T) (both words hold a multiple of kTblRow below
// 65536 in their low half, so v_sad_u16 is |L - R| * kTblRow), table read at
// 2h + that offset (the table starts at LDS address 0).
__device__ __forceinline__ uint32_t rr_score(const uint4& rec, const uint3& r, uint32_t Tb) {
  const uint32_t h = __builtin_popcount(rec.y ^ r.y) + __builtin_popcount(rec.x ^ r.x);
  const uint32_t adb = min(__builtin_amdgcn_sad_u16(rec.z, r.z, 0u), Tb);
  return row_tbl(h, adb);
}

struct RrRaw {        // LDS operands of 8 columns
  uint4 rec[8];
  uint3 rr[8];
};
struct RrBlk {        // their scores and filter words
  uint32_t c[8];
  uint32_t w[8];
};

__device__ __forceinline__ void rr_load_raw(const uint4* lr, const uint4* rrec, RrRaw& R, int x0,
                                            int d) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    R.rec[u] = lr[x0 + u];
    R.rr[u] = *reinterpret_cast<const uint3*>(&rrec[x0 + u - d]);
  }
}

__device__ __forceinline__ void rr_score_raw(const RrRaw& R, RrBlk& S, uint32_t Tb) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    S.c[u] = rr_score(R.rec[u], R.rr[u], Tb);
    S.w[u] = R.rec[u].w;
  }
}

// The same for a block at the row's end: columns from xe on (the right
// border, and the padding past W) score kOff.
__device__ __forceinline__ void rr_score_raw_e(const RrRaw& R, RrBlk& S, uint32_t Tb, int x0,
                                               int xe) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    S.c[u] = x0 + u < xe ? rr_score(R.rec[u], R.rr[u], Tb) : uint32_t(kOff);
    S.w[u] = R.rec[u].w;
  }
}

// The same for a block of the ragged band, where lane d's window starts at
// column lim = hw + d: the right record's index is clamped at 0 and the score
// left of lim is kOff (a zero score).
__device__ __forceinline__ void rr_load_raw_m(const uint4* lr, const uint4* rrec, RrRaw& R,
                                              int x0, int d) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    R.rec[u] = lr[x0 + u];
    R.rr[u] = *reinterpret_cast<const uint3*>(&rrec[max(x0 + u - d, 0)]);
  }
}
__device__ __forceinline__ void rr_score_raw_m(const RrRaw& R, RrBlk& S, uint32_t Tb, int x0,
                                               int lim) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    S.c[u] = x0 + u >= lim ? rr_score(R.rec[u], R.rr[u], Tb) : uint32_t(kOff);
    S.w[u] = R.rec[u].w;
  }
}

// Block n's LDS operands loaded column by column between the scoring steps of
// block c (loaded one step earlier), so the wave's 16 operand reads spread
// over the block's VALU work instead of leaving in one burst ahead of it.
__device__ __forceinline__ void rr_load_score(const uint4* lr, const uint4* rrec, RrRaw& Rn,
                                              int x0n, int d, const RrRaw& Rc, RrBlk& S,
                                              uint32_t Tb) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    Rn.rec[u] = lr[x0n + u];
    Rn.rr[u] = *reinterpret_cast<const uint3*>(&rrec[x0n + u - d]);
    S.c[u] = rr_score(Rc.rec[u], Rc.rr[u], Tb);
    S.w[u] = Rc.rec[u].w;
  }
}

// Forward chain of one aligned block; its values packed two a word (hi16 of
// consecutive steps, low then high) for the block's home. The sixteen
// multiply-adds are one asm statement, each accumulator issued between two
// chain steps (the compiler pads every separate inline-asm result with a
// wait state before its first use).
__device__ __forceinline__ uint32_t rr_chain(const RrBlk& S, uint32_t t2, uint32_t K, uint4& out) {
  uint32_t t[8];
  asm("v_mad_u32_u16 %0, %16, %8, %25\n\t"
      "v_mad_u32_u16 %1, %17, %9, %25\n\t"
      "v_mad_u32_u16 %0, %16, %24, %0 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %2, %18, %10, %25\n\t"
      "v_mad_u32_u16 %1, %17, %0, %1 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %3, %19, %11, %25\n\t"
      "v_mad_u32_u16 %2, %18, %1, %2 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %4, %20, %12, %25\n\t"
      "v_mad_u32_u16 %3, %19, %2, %3 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %5, %21, %13, %25\n\t"
      "v_mad_u32_u16 %4, %20, %3, %4 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %6, %22, %14, %25\n\t"
      "v_mad_u32_u16 %5, %21, %4, %5 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %7, %23, %15, %25\n\t"
      "v_mad_u32_u16 %6, %22, %5, %6 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %7, %23, %6, %7 op_sel:[1,1,0,0]"
      : "=&v"(t[0]), "=&v"(t[1]), "=&v"(t[2]), "=&v"(t[3]), "=&v"(t[4]), "=&v"(t[5]),
        "=&v"(t[6]), "=&v"(t[7])
      : "v"(S.c[0]), "v"(S.c[1]), "v"(S.c[2]), "v"(S.c[3]), "v"(S.c[4]), "v"(S.c[5]),
        "v"(S.c[6]), "v"(S.c[7]), "v"(S.w[0]), "v"(S.w[1]), "v"(S.w[2]), "v"(S.w[3]),
        "v"(S.w[4]), "v"(S.w[5]), "v"(S.w[6]), "v"(S.w[7]), "v"(t2), "v"(K));
  out.x = __builtin_amdgcn_perm(t[1], t[0], 0x07060302u);
  out.y = __builtin_amdgcn_perm(t[3], t[2], 0x07060302u);
  out.z = __builtin_amdgcn_perm(t[5], t[4], 0x07060302u);
  out.w = __builtin_amdgcn_perm(t[7], t[6], 0x07060302u);
  return t[7];
}

// Backward chain over one block, columns 7 down to 0: column j's forward value
// is half j & 1 of packed word j >> 1 (read through op_sel), its coefficient
// word w[j] (of column j + 1); t[j] receives 2 g'(j) + (low bits).
__device__ __forceinline__ void rr_chain_back(const uint32_t (&w)[8], const uint4& v, uint32_t t2,
                                              uint32_t K, uint32_t (&t)[8]) {
  asm("v_mad_u32_u16 %7, %15, %19, %21 op_sel:[0,1,0,0]\n\t"
      "v_mad_u32_u16 %6, %14, %19, %21\n\t"
      "v_mad_u32_u16 %7, %15, %20, %7 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %5, %13, %18, %21 op_sel:[0,1,0,0]\n\t"
      "v_mad_u32_u16 %6, %14, %7, %6 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %4, %12, %18, %21\n\t"
      "v_mad_u32_u16 %5, %13, %6, %5 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %3, %11, %17, %21 op_sel:[0,1,0,0]\n\t"
      "v_mad_u32_u16 %4, %12, %5, %4 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %2, %10, %17, %21\n\t"
      "v_mad_u32_u16 %3, %11, %4, %3 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %1, %9, %16, %21 op_sel:[0,1,0,0]\n\t"
      "v_mad_u32_u16 %2, %10, %3, %2 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %0, %8, %16, %21\n\t"
      "v_mad_u32_u16 %1, %9, %2, %1 op_sel:[1,1,0,0]\n\t"
      "v_mad_u32_u16 %0, %8, %1, %0 op_sel:[1,1,0,0]"
      : "=&v"(t[0]), "=&v"(t[1]), "=&v"(t[2]), "=&v"(t[3]), "=&v"(t[4]), "=&v"(t[5]),
        "=&v"(t[6]), "=&v"(t[7])
      : "v"(w[0]), "v"(w[1]), "v"(w[2]), "v"(w[3]), "v"(w[4]), "v"(w[5]), "v"(w[6]), "v"(w[7]),
        "v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w), "v"(t2), "v"(K));
}

// Filter words of columns x0+1 .. x0+8 (the backward step into column x takes
// the coefficient of x + 1), read before any tile write of the block.
__device__ __forceinline__ void rr_coef8(uint32_t (&w)[8], const uint4* lr, int x0) {
#pragma unroll
  for (int j = 0; j < 8; ++j) w[j] = lr[x0 + j + 1].w;
}

// Final value of column xc (the high half of its t2) straight to the volume:
// lane l writes disparity l of its group, so a wave's store covers 128
// contiguous bytes of the column.
template <uint32_t XB>
__device__ __forceinline__ void rr_put(const RowOut& o, int xc, uint32_t t2) {
  *reinterpret_cast<DE_GLOBAL uint16_t*>(o.vrow + size_t(xc) * XB + 2u * o.lane) =
      uint16_t(t2 >> 16);
}

// H-bwd over P blocks b, b-1, ..., b-P+1 from their packed forward words. The
// next block's coefficients are read while the current block's chain runs,
// and each column's final values leave as one 2-byte store a lane (no LDS
// round trip between the chain and the volume).
template <int P, uint32_t XB>
__device__ __forceinline__ uint32_t rr_back_blocks(const uint4 (&v)[P], uint32_t t2,
                                                   const uint4* lr, int b, const RowOut& o,
                                                   uint32_t K) {
  uint32_t wc[8], wn[8];
  rr_coef8(wc, lr, 8 * b);
#pragma unroll
  for (int k = 0; k < P; ++k) {
    if (k + 1 < P) rr_coef8(wn, lr, 8 * (b - k - 1));
    uint32_t t[8];
    rr_chain_back(wc, v[k], t2, K, t);
    t2 = t[0];
    DE_GLOBAL uint8_t* const p = o.vrow + size_t(8 * (b - k)) * XB + 2u * o.lane;
#pragma unroll
    for (int j = 7; j >= 0; --j)
      *reinterpret_cast<DE_GLOBAL uint16_t*>(p + j * XB) = uint16_t(t[j] >> 16);
#pragma unroll
    for (int j = 0; j < 8; ++j) wc[j] = wn[j];
  }
  return t2;
}

// H-bwd over blocks bhi, bhi-1, ..., blo whose packed words come from ld(b)
// (LDS or the global scratch), P blocks in flight in two alternating sets.
template <int P, uint32_t XB, class LD>
__device__ __forceinline__ uint32_t rr_back_range(const LD& ld, uint32_t t2, const uint4* lr,
                                                  int bhi, int blo, const RowOut& o, uint32_t K) {
  int bb = bhi;
  auto fetch = [&](uint4 (&v)[P], int top) {
#pragma unroll
    for (int k = 0; k < P; ++k) v[k] = ld(top - k);
  };
  if (bb - P + 1 >= blo) {
    uint4 A[P], B[P];
    fetch(A, bb);
    for (; bb - 3 * P + 1 >= blo; bb -= 2 * P) {
      fetch(B, bb - P);
      t2 = rr_back_blocks<P, XB>(A, t2, lr, bb, o, K);
      fetch(A, bb - 2 * P);
      t2 = rr_back_blocks<P, XB>(B, t2, lr, bb - P, o, K);
    }
    if (bb - 2 * P + 1 >= blo) {
      fetch(B, bb - P);
      t2 = rr_back_blocks<P, XB>(A, t2, lr, bb, o, K);
      t2 = rr_back_blocks<P, XB>(B, t2, lr, bb - P, o, K);
      bb -= 2 * P;
    } else {
      t2 = rr_back_blocks<P, XB>(A, t2, lr, bb, o, K);
      bb -= P;
    }
  }
  for (; bb >= blo; --bb) {
    uint4 v[1] = {ld(bb)};
    t2 = rr_back_blocks<1, XB>(v, t2, lr, bb, o, K);
  }
  return t2;
}

// LDS bytes of a row-pass workgroup: table | left records | right records,
// then per wave nl forward slots of 1 KB.
__host__ __device__ inline size_t row_ws_fixed(int W, int T, int G) {
  (void)G;
  const size_t tb = (size_t(T + 1) * kTblRow + 15) & ~size_t(15);
  return tb + 2 * size_t(W + kRowPad) * 16;
}

// G: disparity groups of every frame in the launch; P: H-bwd prefetch depth
// of the global-scratch blocks; nl: LDS forward slots per wave (at least
// kRowTailSlots).
template <int G, int P>
__global__ __launch_bounds__(64 * G) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_row(BatchView b, int base, int nl) {
  constexpr int Dp = G * kWave;
  constexpr uint32_t XB = 2u * Dp;                       // bytes per volume column
  constexpr int NA = kRowABlocks, NV = kRowVBlocks, NT = kRowRegBlocks;
  static_assert(NA % 4 == 0, "AGPR blocks go back in fours");
  extern __shared__ __align__(16) uint8_t smem[];
  const int item = base + blockIdx.x;
  const int f = find_frame(b, kStageRow, item);
  const FrameDesc F = b.frames[f];
  const int y = local_item(b, kStageRow, item, f);
  const int W = F.W, H = F.H;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  const int d = F.dmin + g * kWave + int(lane);
  RowOut o;
  o.vrow = reinterpret_cast<DE_GLOBAL uint8_t*>(gptr(F.vol) + (size_t(y) * W * Dp + g * kWave));
  o.lane = lane;
  o.so = (lane >> 3) * XB + (lane & 7) * 16;
  o.W = W;
  const int xlo = F.hw + F.dmin + g * kWave;             // lane 0 valid from here
  const int xend = W - F.hw;                             // left pixel valid below
  auto zero_row = [&] {
    // every score of this row is zero -> the filtered row is zero (v' = kOff)
    for (int x = 0; x < W; ++x)
      *reinterpret_cast<DE_GLOBAL uint16_t*>(o.vrow + size_t(x) * XB + 2u * lane) = uint16_t(kOff);
  };
  if (y < F.hh || y >= H - F.hh) {                       // uniform over the workgroup
    zero_row();
    return;
  }
  const int nb = (2 * F.hw + 1) * (2 * F.hh + 1) - 1;
  const int T = F.ad_T;
  const int ntt = (T + 1) * int(kTblRow / 2);            // transposed table entries
  const int tb = (ntt * 2 + 15) & ~15;
  uint4* lr = reinterpret_cast<uint4*>(smem + tb);
  uint4* rrec = lr + (W + kRowPad);
  uint8_t* const lf = reinterpret_cast<uint8_t*>(rrec + (W + kRowPad)) + g * (nl * 1024);
  o.tile = nullptr;                                      // stores go straight to the volume
  {
    // entry (a, h) = score of census distance h, intensity difference a <= T;
    // the frame's table holds it at (h << sc_shift) + a
    uint16_t* tbl = reinterpret_cast<uint16_t*>(smem);
    stage_score_table(tbl, cptr(F.sc), ntt, nb, F.sc_shift, threadIdx.x, 64 * G);
    const DE_CONST uint4* lrg = cptr(reinterpret_cast<const uint4*>(F.lrec)) + size_t(y) * W;
    const DE_CONST uint64_t* crg = cptr(F.cr) + size_t(y) * W;
    const DE_CONST uint8_t* Rg = cptr(F.R) + size_t(y) * W;
    // kStage columns a thread in flight: every load of the group is issued
    // before the first LDS store (one memory latency a group, not a column)
    constexpr int nt = 64 * G;
    for (int x0 = threadIdx.x; x0 < W + kRowPad; x0 += kStage * nt) {
      uint4 rec[kStage];
      uint64_t c[kStage];
      uint32_t r[kStage];
#pragma unroll
      for (int u = 0; u < kStage; ++u) {
        const int x = x0 + u * nt;
        const bool in = x < W;
        rec[u] = make_uint4(0u, 0u, 0u, 0u);
        c[u] = 0ull;
        r[u] = 0u;
        if (in) {
          rec[u] = lrg[x];
          c[u] = crg[x];
          r[u] = Rg[x];
        }
      }
#pragma unroll
      for (int u = 0; u < kStage; ++u) {
        const int x = x0 + u * nt;
        if (x < W + kRowPad) {
          lr[x] = rec[u];
          rrec[x] = make_uint4(uint32_t(c[u]), uint32_t(c[u] >> 32), kTblRow * r[u], 0u);
        }
      }
    }
  }
  __syncthreads();
  if (xend <= xlo) {                                     // uniform over the wave
    zero_row();
    return;
  }
  const uint32_t K = 32768u;
  const uint32_t Tb = kTblRow * uint32_t(T);
  const int xa = xlo;
  const int xb = max(xa, min(xlo + kWave - 1, xend));
  const int xm0 = min((xb + 7) & ~7, xend);              // aligned blocks [xm0, 8 B1)
  const int nblk = max(0, ((xend & ~7) - xm0) >> 3);
  const int B0 = xm0 >> 3, B1 = B0 + nblk;
  const int nreg = nblk >= NV ? min(nblk, NT) : 0;       // the last nreg blocks in registers
  const int Bd = B1 - nreg;                              // blocks [B0, Bd) in LDS or global
  const int npre = Bd - B0;
  const int B0p = B0 + (npre & 1);                       // first pipelined block
  // blocks [bL, Bd) (and those from B1 on) live in LDS, lower ones in the
  // global scratch; the pipelined blocks on each side come in pairs
  int bL = max(xa >> 3, Bd - (nl - kRowTailSlots));
  if (bL > B0p && ((bL - B0p) & 1)) ++bL;
  bL = min(bL, Bd);
  const int Wb = (W + 7) >> 3;
  o.frow = reinterpret_cast<DE_GLOBAL uint8_t*>(gptr(F.fwd)) + size_t(y * G + g) * Wb * 1024;
  const uint32_t fo = 16u * lane;                        // lane's bytes within a block
  auto slot = [&](int bk) -> uint32_t {
    return uint32_t(bk >= B1 ? bk - B1 : kRowTailSlots + (Bd - 1 - bk)) * 1024u;
  };
  auto lds_block = [&](int bk) -> uint8_t* { return lf + slot(bk) + fo; };
  auto glb_block = [&](int bk) { return o.frow + size_t(bk) * 1024 + fo; };
  auto store_col = [&](int xc, uint32_t v) {
    const int bk = xc >> 3;
    if (bk >= bL) *reinterpret_cast<uint16_t*>(lds_block(bk) + (xc & 7) * 2) = uint16_t(v);
    else *reinterpret_cast<DE_GLOBAL uint16_t*>(glb_block(bk) + (xc & 7) * 2) = uint16_t(v);
  };
  auto load_col = [&](int xc) -> uint32_t {
    const int bk = xc >> 3;
    if (bk >= bL) return *reinterpret_cast<const uint16_t*>(lds_block(bk) + (xc & 7) * 2);
    return *reinterpret_cast<const DE_GLOBAL uint16_t*>(glb_block(bk) + (xc & 7) * 2);
  };
  auto store_lds = [&](int bk, const uint4& v) {
    *reinterpret_cast<uint4*>(lds_block(bk)) = v;
  };
  auto store_glb = [&](int bk, const uint4& v) {
    *reinterpret_cast<DE_GLOBAL uint4*>(glb_block(bk)) = v;
  };

  // ---- H-fwd. x in [0, xa): every score is zero, so f' stays kOff (not
  // stored: the backward pass uses kOff there) ----
  uint32_t t2 = uint32_t(kOff) << 16;
  int x = xa;
  // The ragged band [xa, xm0), where lane d's window opens at column hw + d,
  // runs the block pipeline from xa's block with masked scores when it ends
  // on a block boundary (every frame but the narrowest): a zero score keeps
  // f' = kOff, so the columns of that block left of xa store kOff, as the
  // backward scan assumes there. Column by column its latency chains cost
  // about 10 us of the row pass at 848x480x128.
  const int Ba = xa >> 3;
  const bool band = (xm0 & 7) == 0 && xm0 > xa;
  RrRaw R[2];
  RrBlk S[2];
  if (band) {
    const int lim = F.hw + d;
    const int Bm = xm0 >> 3;
    auto put = [&](int bk, const uint4& v) {
      if (bk >= bL) store_lds(bk, v);
      else store_glb(bk, v);
    };
    int k = Ba;
    if ((Bm - Ba) & 1) {                                 // one block alone
      uint4 v;
      rr_load_raw_m(lr, rrec, R[0], 8 * k, d);
      rr_score_raw_m(R[0], S[0], Tb, 8 * k, lim);
      t2 = rr_chain(S[0], t2, K, v);
      put(k, v);
      ++k;
    }
    if (k < Bm) {
      // S[0] holds block k scored, R[1] block k+1 loaded; two blocks a step
      rr_load_raw_m(lr, rrec, R[0], 8 * k, d);
      rr_load_raw_m(lr, rrec, R[1], 8 * k + 8, d);
      rr_score_raw_m(R[0], S[0], Tb, 8 * k, lim);
      for (; k < Bm; k += 2) {
        uint4 v;
        rr_load_raw_m(lr, rrec, R[0], 8 * k + 16, d);
        rr_score_raw_m(R[1], S[1], Tb, 8 * k + 8, lim);
        t2 = rr_chain(S[0], t2, K, v);
        put(k, v);
        rr_load_raw_m(lr, rrec, R[1], 8 * k + 24, d);
        rr_score_raw_m(R[0], S[0], Tb, 8 * k + 16, lim);
        t2 = rr_chain(S[1], t2, K, v);
        put(k + 1, v);
      }
    }
    x = xm0;
  } else {
    for (; x < xb; ++x) {                                // ragged band: per-lane window
      const uint4 rec = lr[x];
      const bool ok = x >= F.hw + d;
      const uint3 rv = *reinterpret_cast<const uint3*>(&rrec[ok ? x - d : 0]);
      const uint32_t c = ok ? rr_score(rec, rv, Tb) : uint32_t(kOff);
      t2 = rf_chain_z(rec.w, t2, mad_lo16(rec.w, c, K));
      store_col(x, t2 >> 16);
    }
    for (; x < xm0; ++x) {
      const uint4 rec = lr[x];
      const uint3 rv = *reinterpret_cast<const uint3*>(&rrec[x - d]);
      t2 = rf_chain_z(rec.w, t2, mad_lo16(rec.w, rr_score(rec, rv, Tb), K));
      store_col(x, t2 >> 16);
    }
  }
  if (npre & 1) {                                        // block B0 alone
    uint4 v;
    rr_load_raw(lr, rrec, R[0], x, d);
    rr_score_raw(R[0], S[0], Tb);
    t2 = rr_chain(S[0], t2, K, v);
    if (B0 >= bL) store_lds(B0, v);
    else store_glb(B0, v);
    x += 8;
  }
  const bool piped = Bd - B0p >= 2;
  if (piped) {
    // S[0] holds block k scored, R[1] block k+1 loaded; two blocks a step
    rr_load_raw(lr, rrec, R[0], x, d);
    rr_load_raw(lr, rrec, R[1], x + 8, d);
    rr_score_raw(R[0], S[0], Tb);
    auto pair = [&](int k, auto&& put) {
      uint4 v;
      rr_load_score(lr, rrec, R[0], 8 * k + 16, d, R[1], S[1], Tb);
      t2 = rr_chain(S[0], t2, K, v);
      put(k, v);
      rr_load_score(lr, rrec, R[1], 8 * k + 24, d, R[0], S[0], Tb);
      t2 = rr_chain(S[1], t2, K, v);
      put(k + 1, v);
    };
    const int kg = max(B0p, bL);
    for (int k = B0p; k < kg; k += 2) pair(k, store_glb);
    for (int k = kg; k < Bd; k += 2) pair(k, store_lds);
    x = 8 * Bd;
  }
  uint32_t ag[4 * NA];
  uint4 fr[NV];
  const int j0 = NT - nreg;                              // first register slot in use
  const int Bb = B1 - NT;                                // slot j holds block Bb + j
  if (nreg > 0) {
    if (!piped) {                                        // prime the pipeline at slot j0
      if (j0 & 1) {
        rr_load_raw(lr, rrec, R[1], 8 * Bd, d);
        rr_load_raw(lr, rrec, R[0], 8 * Bd + 8, d);
        rr_score_raw(R[1], S[1], Tb);
      } else {
        rr_load_raw(lr, rrec, R[0], 8 * Bd, d);
        rr_load_raw(lr, rrec, R[1], 8 * Bd + 8, d);
        rr_score_raw(R[0], S[0], Tb);
      }
    }
    auto slot_fwd = [&](auto J) __attribute__((always_inline)) {
      constexpr int j = decltype(J)::value;
      const int xj = 8 * (Bb + j);
      if constexpr (j + 2 < NT) {
        rr_load_score(lr, rrec, R[j & 1], xj + 16, d, R[(j + 1) & 1], S[(j + 1) & 1], Tb);
      } else if constexpr (j + 1 < NT) {
        rr_score_raw(R[(j + 1) & 1], S[(j + 1) & 1], Tb);
      }
      uint4 v;
      t2 = rr_chain(S[j & 1], t2, K, v);
      if constexpr (j < NA) {
        ag[4 * j] = to_acc(v.x);
        ag[4 * j + 1] = to_acc(v.y);
        ag[4 * j + 2] = to_acc(v.z);
        ag[4 * j + 3] = to_acc(v.w);
      } else {
        fr[j - NA] = v;
      }
    };
    // every slot in use (rows of at least NT aligned blocks): straight-line
    // code, no branch between blocks for the pipeline's registers to meet at
    if (j0 == 0) {
      static_for<NT>([&](auto J) __attribute__((always_inline)) { slot_fwd(J); });
    } else {
      static_for<NT>([&](auto J) __attribute__((always_inline)) {
        if (decltype(J)::value >= j0) slot_fwd(J);
      });
    }
    x = 8 * B1;
  }
  if (x == 8 * B1) {
    // the columns from the aligned range's end to W as whole blocks in the
    // tail slots (at most two: xend >= W - 7 and the range ends at xend & ~7),
    // zero scores from xend on; a block may run past W into the records'
    // padding, so the state after column W-1 is its stored value
    for (; x < W; x += 8) {
      uint4 v;
      rr_load_raw_m(lr, rrec, R[0], x, d);                 // (narrow rows: x - d may be < 0)
      rr_score_raw_e(R[0], S[0], Tb, x, xend);
      t2 = rr_chain(S[0], t2, K, v);
      store_lds(x >> 3, v);
    }
    t2 = load_col(W - 1) << 16;
  } else {
    for (; x < xend; ++x) {
      const uint4 rec = lr[x];
      const uint3 rv = *reinterpret_cast<const uint3*>(&rrec[x - d]);
      t2 = rf_chain_z(rec.w, t2, mad_lo16(rec.w, rr_score(rec, rv, Tb), K));
      store_col(x, t2 >> 16);
    }
    for (; x < W; ++x) {                                 // right border: zero scores
      const uint32_t w = lr[x].w;
      t2 = rf_chain_z(w, t2, mad_lo16(w, uint32_t(kOff), K));
      store_col(x, t2 >> 16);
    }
  }

  // ---- H-bwd: g'[W-1] = f'[W-1]; coefficient of x+1 for column x ----
  auto back1 = [&](int xc) {
    const uint32_t fv = xc >= xa ? load_col(xc) : uint32_t(kOff);
    const uint32_t w = lr[xc + 1].w;
    t2 = rf_chain_z(w, t2, mad_lo16(w, fv, K));
    rr_put<XB>(o, xc, t2);
  };
  rr_put<XB>(o, W - 1, t2);
  for (x = W - 2; x >= 8 * B1; --x) back1(x);            // columns from the aligned range's end
  if (nreg > 0) {
    uint4 vr[NV];                                        // block B1 - 1 - k first
#pragma unroll
    for (int k = 0; k < NV; ++k) vr[k] = fr[NV - 1 - k];
    t2 = rr_back_blocks<NV, XB>(vr, t2, lr, B1 - 1, o, K);
    auto chunk_back = [&](auto C) __attribute__((always_inline)) {
      constexpr int c = NA / 4 - 1 - decltype(C)::value;   // chunks from the top down
      uint4 v[4];
      static_for<4>([&](auto Q) __attribute__((always_inline)) {
        constexpr int q = decltype(Q)::value, j = 4 * c + 3 - q;
        v[q] = make_uint4(from_acc(ag[4 * j]), from_acc(ag[4 * j + 1]),
                          from_acc(ag[4 * j + 2]), from_acc(ag[4 * j + 3]));
      });
      t2 = rr_back_blocks<4, XB>(v, t2, lr, Bb + 4 * c + 3, o, K);
    };
    if (j0 == 0) {                                       // straight-line, as in H-fwd
      static_for<NA / 4>([&](auto C) __attribute__((always_inline)) { chunk_back(C); });
    } else {
      static_for<NA / 4>([&](auto C) __attribute__((always_inline)) {
        constexpr int c = NA / 4 - 1 - decltype(C)::value;
        if (4 * c >= j0) {
          chunk_back(C);
        } else if (4 * c + 3 >= j0) {
          static_for<4>([&](auto Q) __attribute__((always_inline)) {
            constexpr int j = 4 * c + 3 - decltype(Q)::value;
            if (j >= j0) {
              uint4 v[1] = {make_uint4(from_acc(ag[4 * j]), from_acc(ag[4 * j + 1]),
                                       from_acc(ag[4 * j + 2]), from_acc(ag[4 * j + 3]))};
              t2 = rr_back_blocks<1, XB>(v, t2, lr, Bb + j, o, K);
            }
          });
        }
      });
    }
  }
  // blocks below the registers: LDS [max(blo, bL), Bd), then global [blo, bL)
  const int blo = band ? Ba : min((xa + 7) >> 3, Bd);
  t2 = rr_back_range<4, XB>(
      [&](int bk) { return *reinterpret_cast<const uint4*>(lds_block(bk)); }, t2, lr, Bd - 1,
      max(blo, bL), o, K);
  t2 = rr_back_range<P, XB>(
      [&](int bk) { return *reinterpret_cast<const DE_GLOBAL uint4*>(glb_block(bk)); }, t2, lr,
      min(bL, Bd) - 1, blo, o, K);
  if (band) {
    // blocks left of xa's: every forward value is kOff
    const uint32_t ko = uint32_t(kOff) * 0x10001u;
    t2 = rr_back_range<4, XB>([&](int) { return make_uint4(ko, ko, ko, ko); }, t2, lr, Ba - 1, 0,
                              o, K);
  } else {
    for (x = 8 * blo - 1; x >= 0; --x) back1(x);
  }
}

// ---- the streaming row pass (k_rows) ----
//
// One wave per (row, 64-disparity group), each staging its own row, about 210
// VGPRs, so two waves fit a SIMD when a launch (or frames in flight) has more
// waves than SIMDs. The forward values go to a scratch private to the (row,
// group) wave, in blocks of 8 columns laid out [x / 8][lane][x % 8], except
// the last kRowSRegBlocks aligned blocks, which stay in VGPRs; the backward
// pass reads them back P blocks ahead. The record words stay packed as z = kTblRow*L | 2a << 16,
// w = 65536 - 2a (the chain reads 2a through op_sel, the intensity difference
// masks z's low half).

__device__ __forceinline__ void rows_load_raw(const RowLds& q, RowRaw& R, int x0, int d) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    R.rec[u] = q.lr[x0 + u];
    R.cr[u] = q.crs[x0 + u - d];
    R.r2[u] = q.r2s[x0 + u - d];
  }
}

// The ragged band's form of the two above: right index clamped at 0, score
// kOff left of lane d's window (column lim = hw + d).
__device__ __forceinline__ void rows_load_raw_m(const RowLds& q, RowRaw& R, int x0, int d) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int j = max(x0 + u - d, 0);
    R.rec[u] = q.lr[x0 + u];
    R.cr[u] = q.crs[j];
    R.r2[u] = q.r2s[j];
  }
}
__device__ __forceinline__ void row_score_raw_m(const RowLds& q, const RowRaw& R, RowBlk& S,
                                                int x0, int lim) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    S.c[u] = x0 + u >= lim ? row_score(q, R.rec[u], R.cr[u], R.r2[u]) : uint32_t(kOff);
    S.z[u] = R.rec[u].z;
    S.w[u] = R.rec[u].w;
  }
}

// Chain of one aligned 8-column block; its forward values leave as one 16-byte
// word per lane (hi16 of consecutive steps packed low, high).
__device__ __forceinline__ uint32_t rows_chain_fwd(const RowBlk& S, uint32_t t2,
                                                  DE_GLOBAL uint8_t* fw) {
  uint32_t t[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    t2 = rf_chain_z(S.z[u], t2, __umul24(S.w[u], S.c[u]) + 32768u);
    t[u] = t2;
  }
  uint4 v;
  v.x = __builtin_amdgcn_perm(t[1], t[0], 0x07060302u);
  v.y = __builtin_amdgcn_perm(t[3], t[2], 0x07060302u);
  v.z = __builtin_amdgcn_perm(t[5], t[4], 0x07060302u);
  v.w = __builtin_amdgcn_perm(t[7], t[6], 0x07060302u);
  *reinterpret_cast<DE_GLOBAL uint4*>(fw) = v;
  return t2;
}

// The same chain into a register block instead of the forward scratch.
__device__ __forceinline__ uint32_t rows_chain_reg(const RowBlk& S, uint32_t t2, uint4& out) {
  uint32_t t[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    t2 = rf_chain_z(S.z[u], t2, __umul24(S.w[u], S.c[u]) + 32768u);
    t[u] = t2;
  }
  out.x = __builtin_amdgcn_perm(t[1], t[0], 0x07060302u);
  out.y = __builtin_amdgcn_perm(t[3], t[2], 0x07060302u);
  out.z = __builtin_amdgcn_perm(t[5], t[4], 0x07060302u);
  out.w = __builtin_amdgcn_perm(t[7], t[6], 0x07060302u);
  return t2;
}

// Aligned forward blocks a row pass keeps in registers (64 VGPRs): the last
// kRowSRegBlocks of its pipelined blocks skip the forward scratch, and the
// backward scan takes them first.
constexpr int kRowSRegBlocks = 16;

// H-bwd over P interior blocks b, b-1, ..., b-P+1 from their loaded forward
// words, coefficient of column x+1 for column x. The next block's
// coefficients are read before the current block's tile writes; the tiles
// alternate, and a block's 16-byte stores leave after the next block's chain,
// so neither LDS read is waited on right where it is issued.
template <int P, uint32_t XB>
__device__ __forceinline__ uint32_t rows_back_blocks(const uint4 (&v)[P], uint32_t t2,
                                                    const uint4* lr, int b, const RowOut& o) {
  uint2 zc[8], zn[8];
  uint4 pend = make_uint4(0u, 0u, 0u, 0u);
  row_coef8(zc, lr, 8 * b);
#pragma unroll
  for (int k = 0; k < P; ++k) {
    if (k + 1 < P) row_coef8(zn, lr, 8 * (b - k - 1));
    uint16_t* const tile = o.tile + (k & 1) * 8 * kWave;
    const uint32_t wv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
    for (int j = 7; j >= 0; --j) {
      const uint32_t f = (j & 1) ? (wv[j >> 1] >> 16) : (wv[j >> 1] & 0xFFFFu);
      t2 = rf_chain_z(zc[j].x, t2, __umul24(zc[j].y, f) + 32768u);
      tile[j * kWave + o.lane] = uint16_t(t2 >> 16);
    }
    if (k > 0) row_store8<XB>(o, b - k + 1, pend);
    pend = *reinterpret_cast<const uint4*>(tile + o.lane * 8);
#pragma unroll
    for (int j = 0; j < 8; ++j) zc[j] = zn[j];
  }
  row_store8<XB>(o, b - P + 1, pend);
  return t2;
}

// G: disparity groups of every frame in the launch; P: H-bwd prefetch depth in
// 8-column blocks. A workgroup is 1..4 independent waves (one item each, own
// LDS region of pw bytes); the launcher groups waves only when every wave
// fits in one round, where one wave per SIMD of a CU measured faster than
// lone-wave workgroups.
template <int G, int P>
__global__ __launch_bounds__(256) void k_rows(BatchView b, int base, int items, int pw) {
  constexpr int Dp = G * kWave;
  constexpr uint32_t XB = 2u * Dp;                       // bytes per volume column
  extern __shared__ __align__(16) uint8_t smem_all[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nwg = blockDim.x >> 6;
  if (blockIdx.x * nwg + wv >= items) return;            // whole wave: no barriers follow
  uint8_t* const smem = smem_all + wv * pw;
  const int witem = blockIdx.x * nwg + wv;               // wave = (row item, group)
  const int item = base + witem / G, g = witem - (witem / G) * G;
  const int f = find_frame(b, kStageRow, item);
  const FrameDesc F = b.frames[f];
  const int y = local_item(b, kStageRow, item, f);
  const int W = F.W, H = F.H;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t vo = 2u * lane;                         // lane's byte offset in a column
  const int d = F.dmin + g * kWave + int(lane);
  RowOut o;
  o.vrow = reinterpret_cast<DE_GLOBAL uint8_t*>(gptr(F.vol) + (size_t(y) * W * Dp + g * kWave));
  o.lane = lane;
  o.so = (lane >> 3) * XB + (lane & 7) * 16;
  o.W = W;
  const int xlo = F.hw + F.dmin + g * kWave;             // lane 0 valid from here
  const int xend = W - F.hw;                             // left pixel valid below
  if (y < F.hh || y >= H - F.hh || xend <= xlo) {
    // every score of this row is zero -> the filtered row is zero (v' = kOff)
    for (int x = 0; x < W; ++x)
      *reinterpret_cast<DE_GLOBAL uint16_t*>(o.vrow + size_t(x) * XB + vo) = uint16_t(kOff);
    return;
  }
  const int nb = (2 * F.hw + 1) * (2 * F.hh + 1) - 1;
  const int T = F.ad_T;
  const int ntt = (T + 1) * int(kTblRow / 2);            // transposed table entries
  // LDS: score table (transposed) | two transpose tiles (8 x 64 u16) | left
  // records (uint4) | right census (u64) | kTblRow*R (u32)
  o.tile = reinterpret_cast<uint16_t*>(smem + ((ntt * 2 + 15) & ~15));
  uint4* lr = reinterpret_cast<uint4*>(o.tile + 2 * 8 * kWave);
  uint64_t* crs = reinterpret_cast<uint64_t*>(lr + W);
  uint32_t* r2s = reinterpret_cast<uint32_t*>(crs + W);
  {
    // entry (a, h) = score of census distance h, intensity difference a <= T;
    // the frame's table holds it at (h << sc_shift) + a
    uint16_t* tbl = reinterpret_cast<uint16_t*>(smem);
    stage_score_table(tbl, cptr(F.sc), ntt, nb, F.sc_shift, lane, kWave);
    const DE_CONST uint4* lrg = cptr(reinterpret_cast<const uint4*>(F.lrec)) + size_t(y) * W;
    const DE_CONST uint64_t* crg = cptr(F.cr) + size_t(y) * W;
    const DE_CONST uint8_t* Rg = cptr(F.R) + size_t(y) * W;
#pragma unroll 8
    for (int x = lane; x < W; x += kWave) {
      uint4 rec = lrg[x];
      // the packed form this pass reads: kTblRow * L | 2a << 16, 65536 - 2a
      rec.z = rec.z | (rec.w & 0xFFFF0000u);
      rec.w = 65536u - (rec.w >> 16);
      lr[x] = rec;
      crs[x] = crg[x];
      r2s[x] = kTblRow * Rg[x];
    }
  }
  // the wave reads back only what it wrote; LDS keeps one wave's accesses in
  // order, so a compiler barrier suffices
  __builtin_amdgcn_wave_barrier();
  RowLds q;
  q.hb = uint32_t(wv * pw) >> 1;
  q.lr = lr;
  q.crs = crs;
  q.r2s = r2s;
  q.Tb = kTblRow * uint32_t(T);
  const int Wb = (W + 7) >> 3;
  o.frow = reinterpret_cast<DE_GLOBAL uint8_t*>(gptr(F.fwd)) + size_t(y * G + g) * Wb * 1024;
  const uint32_t fo = 16u * lane;                        // lane's bytes within a block
  auto fwd_at = [&](int x) {
    return reinterpret_cast<DE_GLOBAL uint16_t*>(o.frow + size_t(x >> 3) * 1024 + fo +
                                                 (x & 7) * 2);
  };
  const int xa = xlo;                                    // xlo < xend here
  const int xb = max(xa, min(xlo + kWave - 1, xend));

  // ---- H-fwd. x in [0, xa): every score is zero, so f' stays kOff (not
  // stored: the backward pass uses kOff there) ----
  uint32_t t2 = uint32_t(kOff) << 16;
  const int xm0 = min((xb + 7) & ~7, xend);              // aligned blocks [xm0, xm0 + 8 nblk)
  // the ragged band [xa, xm0) as masked blocks from xa's block when it ends on
  // a block boundary (as in k_row: the block's columns left of xa store kOff)
  const int Ba = xa >> 3;
  const bool band = (xm0 & 7) == 0 && xm0 > xa;
  if (band) {
    const int lim = F.hw + d;
    const int Bm = xm0 >> 3;
    RowRaw R0, R1;
    RowBlk S0, S1;
    int k = Ba;
    if ((Bm - Ba) & 1) {                                 // one block alone
      rows_load_raw_m(q, R0, 8 * k, d);
      row_score_raw_m(q, R0, S0, 8 * k, lim);
      t2 = rows_chain_fwd(S0, t2, o.frow + size_t(k) * 1024 + fo);
      ++k;
    }
    if (k < Bm) {
      rows_load_raw_m(q, R0, 8 * k, d);
      rows_load_raw_m(q, R1, 8 * k + 8, d);
      row_score_raw_m(q, R0, S0, 8 * k, lim);
      for (; k < Bm; k += 2) {
        rows_load_raw_m(q, R0, 8 * k + 16, d);
        row_score_raw_m(q, R1, S1, 8 * k + 8, lim);
        t2 = rows_chain_fwd(S0, t2, o.frow + size_t(k) * 1024 + fo);
        rows_load_raw_m(q, R1, 8 * k + 24, d);
        row_score_raw_m(q, R0, S0, 8 * k + 16, lim);
        t2 = rows_chain_fwd(S1, t2, o.frow + size_t(k + 1) * 1024 + fo);
      }
    }
  } else {
    for (int x = xa; x < xb; ++x) {                      // ragged band: per-lane window
      const uint4 rec = lr[x];
      const bool ok = x >= F.hw + d;
      const int j = ok ? x - d : 0;
      const uint32_t c = ok ? row_score(q, rec, crs[j], r2s[j]) : uint32_t(kOff);
      t2 = rf_chain_z(rec.z, t2, __umul24(rec.w, c) + 32768u);
      *fwd_at(x) = uint16_t(t2 >> 16);
    }
  }
  const int nblk_all = max(0, ((xend & ~7) - xm0) >> 3);
  // the last kRowSRegBlocks aligned blocks [E - NB, E) go to registers when at
  // least two blocks precede them
  constexpr int NB = kRowSRegBlocks;
  const bool use_reg = nblk_all >= NB + 2;
  const int nblk = use_reg ? nblk_all - NB : nblk_all;   // blocks through the scratch
  const int E = (xm0 >> 3) + nblk_all;
  uint4 fr[NB];
  int x = band ? xm0 : xb;
  for (; x < xm0; ++x) {
    const uint4 rec = lr[x];
    t2 = rf_chain_z(rec.z, t2,
                    __umul24(rec.w, row_score(q, rec, crs[x - d], r2s[x - d])) + 32768u);
    *fwd_at(x) = uint16_t(t2 >> 16);
  }
  if (nblk >= 2) {
    // block k+2 loaded, k+1 scored, k chained; two register sets of each
    RowRaw R0, R1;
    RowBlk S0, S1;
    rows_load_raw(q, R0, x, d);
    rows_load_raw(q, R1, x + 8, d);
    row_score_raw(q, R0, S0);
    int k = 0;                    // S0 holds block k scored, R1 block k+1 loaded
    for (; k + 4 <= nblk; k += 2, x += 16) {
      rows_load_raw(q, R0, x + 16, d);
      row_score_raw(q, R1, S1);
      t2 = rows_chain_fwd(S0, t2, o.frow + size_t(x >> 3) * 1024 + fo);
      rows_load_raw(q, R1, x + 24, d);
      row_score_raw(q, R0, S0);
      t2 = rows_chain_fwd(S1, t2, o.frow + size_t((x + 8) >> 3) * 1024 + fo);
    }
    if (nblk - k == 3) {
      rows_load_raw(q, R0, x + 16, d);
      row_score_raw(q, R1, S1);
      t2 = rows_chain_fwd(S0, t2, o.frow + size_t(x >> 3) * 1024 + fo);
      row_score_raw(q, R0, S0);
      t2 = rows_chain_fwd(S1, t2, o.frow + size_t((x + 8) >> 3) * 1024 + fo);
      t2 = rows_chain_fwd(S0, t2, o.frow + size_t((x + 16) >> 3) * 1024 + fo);
      x += 24;
    } else {
      row_score_raw(q, R1, S1);
      t2 = rows_chain_fwd(S0, t2, o.frow + size_t(x >> 3) * 1024 + fo);
      t2 = rows_chain_fwd(S1, t2, o.frow + size_t((x + 8) >> 3) * 1024 + fo);
      x += 16;
    }
  } else if (nblk == 1) {
    RowRaw R0;
    RowBlk S0;
    rows_load_raw(q, R0, x, d);
    row_score_raw(q, R0, S0);
    t2 = rows_chain_fwd(S0, t2, o.frow + size_t(x >> 3) * 1024 + fo);
    x += 8;
  }
  if (use_reg) {
    // the same pipeline, unrolled, into fr[0 .. NB): S[j & 1] holds block j
    // scored, R[(j + 1) & 1] block j + 1 loaded
    RowRaw R[2];
    RowBlk S[2];
    rows_load_raw(q, R[0], x, d);
    rows_load_raw(q, R[1], x + 8, d);
    row_score_raw(q, R[0], S[0]);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j + 2 < NB) rows_load_raw(q, R[j & 1], x + 8 * (j + 2), d);
      if (j + 1 < NB) row_score_raw(q, R[(j + 1) & 1], S[(j + 1) & 1]);
      t2 = rows_chain_reg(S[j & 1], t2, fr[j]);
    }
    x += 8 * NB;
  }
  for (; x < xend; ++x) {
    const uint4 rec = lr[x];
    t2 = rf_chain_z(rec.z, t2,
                    __umul24(rec.w, row_score(q, rec, crs[x - d], r2s[x - d])) + 32768u);
    *fwd_at(x) = uint16_t(t2 >> 16);
  }
  for (; x < W; ++x) {                                   // right border: zero scores
    const uint4 rec = lr[x];
    t2 = rf_chain_z(rec.z, t2, __umul24(rec.w, uint32_t(kOff)) + 32768u);
    *fwd_at(x) = uint16_t(t2 >> 16);
  }

  // ---- H-bwd: g'[W-1] = f'[W-1]; coefficient of x+1 for column x ----
  // top block bhi (holds W-1), interior blocks [blo, bhi) from the wide
  // prefetch, then the columns below 8*blo (forward value kOff below xa)
  auto back1 = [&](int xc) {
    const uint32_t fv = xc >= xa ? uint32_t(*fwd_at(xc)) : uint32_t(kOff);
    const uint2 zw = *reinterpret_cast<const uint2*>(&lr[xc + 1].z);
    t2 = rf_chain_z(zw.x, t2, __umul24(zw.y, fv) + 32768u);
    o.tile[(xc & 7) * kWave + lane] = uint16_t(t2 >> 16);
    if ((xc & 7) == 0) row_flush<XB, true>(o, xc >> 3);
  };
  const int bhi = (W - 1) >> 3;
  const int blo = min(bhi, band ? Ba : (xa + 7) >> 3);
  o.tile[((W - 1) & 7) * kWave + lane] = uint16_t(t2 >> 16);
  if (((W - 1) & 7) == 0) row_flush<XB, true>(o, bhi);
  for (x = W - 2; x >= 8 * bhi; --x) back1(x);
  int bb = bhi - 1;                                      // next interior block
  if (use_reg)
    for (; bb >= E; --bb) {                              // at most one block above E
      uint4 v[1];
      row_load_blocks<1>(v, o, bb);
      t2 = rows_back_blocks<1, XB>(v, t2, lr, bb, o);
    }
  if (use_reg) bb = E - NB - 1;                          // below the register blocks
  // two register sets alternate between loading and filtering
  uint4 A[P], B[P];
  const bool grouped = bb - P + 1 >= blo;
  if (grouped) row_load_blocks<P>(A, o, bb);             // in flight over the register blocks
  if (use_reg) {
    uint4 vr[NB];                                        // block E - 1 - k first
#pragma unroll
    for (int k = 0; k < NB; ++k) vr[k] = fr[NB - 1 - k];
    t2 = rows_back_blocks<NB, XB>(vr, t2, lr, E - 1, o);
  }
  if (grouped) {
    for (; bb - 3 * P + 1 >= blo; bb -= 2 * P) {
      row_load_blocks<P>(B, o, bb - P);
      t2 = rows_back_blocks<P, XB>(A, t2, lr, bb, o);
      row_load_blocks<P>(A, o, bb - 2 * P);
      t2 = rows_back_blocks<P, XB>(B, t2, lr, bb - P, o);
    }
    if (bb - 2 * P + 1 >= blo) {
      row_load_blocks<P>(B, o, bb - P);
      t2 = rows_back_blocks<P, XB>(A, t2, lr, bb, o);
      t2 = rows_back_blocks<P, XB>(B, t2, lr, bb - P, o);
      bb -= 2 * P;
    } else {
      t2 = rows_back_blocks<P, XB>(A, t2, lr, bb, o);
      bb -= P;
    }
  }
  for (; bb >= blo; --bb) {
    uint4 v[1];
    row_load_blocks<1>(v, o, bb);
    t2 = rows_back_blocks<1, XB>(v, t2, lr, bb, o);
  }
  if (band) {
    // blocks left of xa's: every forward value is kOff
    const uint32_t ko = uint32_t(kOff) * 0x10001u;
    for (int bk = Ba - 1; bk >= 0; --bk) {
      uint4 v[1] = {make_uint4(ko, ko, ko, ko)};
      t2 = rows_back_blocks<1, XB>(v, t2, lr, bk, o);
    }
  } else {
    for (x = 8 * blo - 1; x >= 0; --x) back1(x);
  }
}

// ---------------------------------------------------------------------------
// Host launchers.

// Launch variants (chosen per shape bucket by de_autotune).
// row pass: resident (global-scratch prefetch 16 blocks, forward slots in LDS
// or not), streaming (backward prefetch 8, 16 blocks)
const int kRowVariants[kNumRowVariants] = {16, 16, 8, 16};

int device_simds() {
  int lds = 0, cus = 1;
  device_shape(&lds, &cus);
  return 4 * cus;
}

// Smallest LDS a row-pass workgroup can run with (kRowTailSlots slots a wave).
size_t row_pass_min_lds(int W, int T, int G) {
  return row_ws_fixed(W, T, G) + size_t(kRowTailSlots) * G * 1024;
}

template <int G, int P>
static hipError_t launch_row_gp(const BatchView& b, int base, int items, int maxW, int T,
                                bool lds_tier, hipStream_t s) {
  int lds = 0, cus = 1;
  device_shape(&lds, &cus);
  static const hipError_t attr = hipFuncSetAttribute(
      reinterpret_cast<const void*>(&k_row<G, P>), hipFuncAttributeMaxDynamicSharedMemorySize,
      lds);
  if (attr != hipSuccess) return attr;
  // one wave per SIMD (a wave holds up to 512 registers): 4 / G workgroups a
  // CU while their LDS fits, each taking an equal share of it; the share left
  // after the staged row goes to forward slots
  const size_t fixed = row_ws_fixed(maxW, T, G);
  const size_t per_slot = size_t(G) * 1024;
  if (fixed + kRowTailSlots * per_slot > size_t(lds)) return hipErrorInvalidValue;
  int wgpc = std::max(1, 4 / G);
  while (wgpc > 1 && fixed + kRowTailSlots * per_slot > (size_t(lds) / wgpc & ~size_t(511)))
    --wgpc;
  const size_t budget = size_t(lds) / wgpc & ~size_t(511);
  const int nl = lds_tier ? int(std::min<size_t>(512, (budget - fixed) / per_slot))
                          : kRowTailSlots;
  hipLaunchKernelGGL((k_row<G, P>), dim3(items), dim3(64 * G), fixed + size_t(nl) * per_slot, s,
                     b, base, nl);
  return hipGetLastError();
}

template <int G, int P>
static hipError_t launch_rows_gp(const BatchView& b, int base, int items, size_t pw,
                                hipStream_t s) {
  int lds = 0, cus = 1;
  device_shape(&lds, &cus);
  static const hipError_t attr = hipFuncSetAttribute(
      reinterpret_cast<const void*>(&k_rows<G, P>), hipFuncAttributeMaxDynamicSharedMemorySize,
      lds);
  if (attr != hipSuccess) return attr;
  pw = (pw + 15) & ~size_t(15);
  // waves per workgroup: when every wave fits in one round (at most one per
  // SIMD), as many per workgroup as spread them one per SIMD of a CU (and the
  // LDS allows); with more waves than SIMDs, lone-wave workgroups spread
  // more evenly and their waves overlap on a SIMD
  const int fit = int(std::min<size_t>(4, size_t(lds) / pw));
  const int nw = items <= 4 * cus ? std::max(1, std::min(fit, (items + cus - 1) / cus)) : 1;
  hipLaunchKernelGGL((k_rows<G, P>), dim3((items + nw - 1) / nw), dim3(64 * nw), nw * pw, s, b,
                     base, items, int(pw));
  return hipGetLastError();
}

// Variants: 0, 1 the register-resident pass with and without forward slots in
// LDS (global-scratch prefetch 16 blocks), 2, 3 the streaming pass (backward
// prefetch 8 or 16 blocks).
template <int G>
static hipError_t launch_row_g(const BatchView& b, int base, int items, int maxW, int T,
                               size_t lds_wave, int variant, hipStream_t s) {
  switch (variant) {
    case 0: return launch_row_gp<G, 16>(b, base, items, maxW, T, true, s);
    case 1: return launch_row_gp<G, 16>(b, base, items, maxW, T, false, s);
    case 2: return launch_rows_gp<G, 8>(b, base, items * G, lds_wave, s);
    default: return launch_rows_gp<G, 16>(b, base, items * G, lds_wave, s);
  }
}

// Row pass over the frames whose rows are work items [base, base + items);
// every one of them must have exactly G disparity groups. maxW: the widest of
// them; T: their intensity truncation (rows of the staged score table);
// lds_wave: LDS a streaming-pass wave stages.
hipError_t launch_row(const BatchView& b, int base, int items, int G, int maxW, int T,
                      size_t lds_wave, int variant, hipStream_t s) {
  if (variant < 0 || variant >= kNumRowVariants) return hipErrorInvalidValue;
  if (items <= 0) return hipSuccess;
  switch (G) {
    case 1: return launch_row_g<1>(b, base, items, maxW, T, lds_wave, variant, s);
    case 2: return launch_row_g<2>(b, base, items, maxW, T, lds_wave, variant, s);
    case 3: return launch_row_g<3>(b, base, items, maxW, T, lds_wave, variant, s);
    case 4: return launch_row_g<4>(b, base, items, maxW, T, lds_wave, variant, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace de
