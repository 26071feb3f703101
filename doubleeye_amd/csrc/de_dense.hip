// Dense image-plane stages for gfx950: census + filter coefficients, the fused
// score + horizontal recursive filter (row pass), and the fused vertical
// recursive filter + exact top-2 (column pass).
//
// Reference semantics (paths relative to the reference repository root):
//   census          core/tools/de_dense.cpp:94-129, core/src/preproc.cpp:110-138
//   coefficients    core/tools/de_dense.cpp:531-561
//   graded score    core/tools/de_dense.cpp:1201-1243
//   int16 filter    core/tools/de_dense.cpp:411-430, 485-508 (whole plane, order
//                   H-fwd, H-bwd, V-fwd, V-bwd)
//   top-2 insert    core/tools/de_dense.cpp:1249-1294 (ascending k, strictly
//                   greater displaces) with the sub-pixel neighbour capture
// Every integer result is bit-identical to the reference; see de_common.h for
// the offset representation that makes one filter step two multiply-adds.
//
// Volume layout: v' u16 at vol[(y*W + x)*Dpad + k], Dpad = 64*G. A wave owns 64
// consecutive k: the column pass reads and writes one 128-byte segment per row,
// the row pass writes 8 columns at a time with 16-byte stores.
//
// The recurrences are sequential along their scan line, so a wave runs 64 (row
// pass) or 64*G (column pass) independent chains; latency is hidden by
// prefetching the scan U steps ahead, not by occupancy.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <utility>

#include "de_batch.h"
#include "de_common.h"

namespace de {

// Row stride in bytes of the score table as the row passes stage it (66 u16:
// census distances 0..62 and padding).
constexpr uint32_t kTblRow = 132;

// 2*f' = a2 * hi16(t2prev) + acc, with a2 = 2a wave-uniform (SGPR) and
// acc = (65536 - a2) * c' + 32768 computed off the dependency chain.
__device__ __forceinline__ uint32_t rf_chain(uint32_t a2, uint32_t t2prev, uint32_t acc) {
  uint32_t r;
  asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[0,1,0,0]" : "=v"(r) : "s"(a2), "v"(t2prev), "v"(acc));
  return r;
}
__device__ __forceinline__ uint32_t rf_chain_v(uint32_t a2, uint32_t t2prev, uint32_t acc) {
  uint32_t r;
  asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[0,1,0,0]" : "=v"(r) : "v"(a2), "v"(t2prev), "v"(acc));
  return r;
}
__device__ __forceinline__ uint32_t rf_acc(uint32_t a2, uint32_t c) {
  return __umul24(65536u - a2, c) + 32768u;
}
__device__ __forceinline__ uint16_t hi16(uint32_t t2) { return uint16_t(t2 >> 16); }

__device__ __forceinline__ uint32_t sad_u32(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_sad_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

__device__ __forceinline__ uint32_t med3(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

__device__ __forceinline__ int local_item(const BatchView& b, int stage, int item, int f) {
  return item - b.start[stage * (b.n + 1) + f];
}

// Census of the pixel at (cy, cx) of an LDS tile: bit b is set when window
// position b (row-major over (2HW+1) x (2HH+1), the centre skipped) is darker
// than the centre. The bits enter from the last position down: n - c (bytes,
// so bit 31 is n < c) is funnel-shifted into the word, acc = (acc << 1) |
// ((n - c) >> 31), two instructions a bit and no lane masks.
template <int HW, int HH, int TW>
__device__ __forceinline__ uint64_t census_at(const uint8_t (*t)[TW], int cy, int cx) {
  constexpr int NWX = 2 * HW + 1, NB = NWX * (2 * HH + 1) - 1, C = NB / 2;
  const uint32_t c = t[cy][cx];
  uint32_t hi = 0, lo = 0;
#pragma unroll
  for (int b = NB - 1; b >= 0; --b) {
    const int p = b < C ? b : b + 1;
    const uint32_t diff = uint32_t(t[cy + p / NWX - HH][cx + p % NWX - HW]) - c;
    if (b >= 32) hi = __builtin_amdgcn_alignbit(hi, diff, 31);
    else lo = __builtin_amdgcn_alignbit(lo, diff, 31);
  }
  return uint64_t(lo) | (uint64_t(hi) << 32);
}

// ---------------------------------------------------------------------------
// Stage 1: census of both images, left record, transposed vertical coefficients.
// Tile of kPrepTileW x kPrepTileH pixels of both images staged in LDS with the
// census halo, rows padded to a multiple of 4 bytes. 256 threads: threads
// 0..127 take the left image, 128..255 the right, each 4 horizontally
// adjacent pixels of one tile row. A thread reads each window row's 12 bytes
// (its 4 pixels' windows together, 4-byte aligned) as three words, and each
// census bit is one byte-selecting subtract (n - c: bit 31 set when n < c)
// funnel-shifted into the word: two instructions a bit, two LDS reads a window
// row for 4 pixels.
template <int HW, int HH>
__global__ __launch_bounds__(256) void k_prep(BatchView b) {
  static_assert(2 * HW + 4 <= 12, "a window row of 4 pixels spans three words");
  constexpr int NWX = 2 * HW + 1, NB = NWX * (2 * HH + 1) - 1, C = NB / 2;
  constexpr int TWP = (kPrepTileW + 2 * HW + 3) & ~3;     // padded tile row (bytes)
  constexpr int TH = kPrepTileH + 2 * HH;
  static_assert(kPrepTileW * kPrepTileH == 4 * 128, "4 pixels a thread, 128 threads an image");
  __shared__ __align__(16) uint8_t st[2][TH][TWP];
  __shared__ uint16_t srf[256];
  const int item = blockIdx.x;
  const int f = find_frame(b, kStagePrep, item);
  const FrameDesc F = b.frames[f];
  const int local = local_item(b, kStagePrep, item, f);
  const int W = F.W, H = F.H;
  const int ntx = (W + kPrepTileW - 1) / kPrepTileW;
  const int x0 = (local % ntx) * kPrepTileW, y0 = (local / ntx) * kPrepTileH;
  const DE_CONST uint8_t* Lg = cptr(F.L);
  const DE_CONST uint8_t* Rg = cptr(F.R);
  srf[threadIdx.x] = cptr(F.rf)[threadIdx.x];
  for (int i = threadIdx.x; i < TWP * TH; i += 256) {
    const int ty = i / TWP, tx = i - ty * TWP;
    const int gx = x0 - HW + tx, gy = y0 - HH + ty;
    const bool in = gx >= 0 && gx < W && gy >= 0 && gy < H;
    st[0][ty][tx] = in ? Lg[size_t(gy) * W + gx] : 0;
    st[1][ty][tx] = in ? Rg[size_t(gy) * W + gx] : 0;
  }
  __syncthreads();
  const int img = threadIdx.x >> 7;                      // wave-uniform
  const int u = threadIdx.x & 127;
  const int r = u >> 4, xq = (u & 15) * 4;               // tile row, first of 4 columns
  const int y = y0 + r;
  if (y >= H) return;
  // window rows dy = -HH..HH of the 4 pixels: bytes xq .. xq+11 of tile row r+HH+dy
  uint32_t wr[2 * HH + 1][3];
#pragma unroll
  for (int dy = 0; dy < 2 * HH + 1; ++dy) {
    const uint32_t* rowp = reinterpret_cast<const uint32_t*>(&st[img][r + dy][xq]);
    wr[dy][0] = rowp[0];
    wr[dy][1] = rowp[1];
    wr[dy][2] = rowp[2];
  }
  auto byte_at = [&](int dy, int k) -> uint32_t {   // byte k of window row dy (constants)
    return (wr[dy][k >> 2] >> (8 * (k & 3))) & 0xFFu;
  };
  const bool rows_ok = y >= HH && y < H - HH;
  uint64_t cen[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int x = x0 + xq + i;
    const uint32_t c = byte_at(HH, i + HW);
    uint32_t hi = 0, lo = 0;
#pragma unroll
    for (int bb = NB - 1; bb >= 0; --bb) {
      const int p = bb < C ? bb : bb + 1;
      const uint32_t diff = byte_at(p / NWX, i + p % NWX) - c;
      if (bb >= 32) hi = __builtin_amdgcn_alignbit(hi, diff, 31);
      else lo = __builtin_amdgcn_alignbit(lo, diff, 31);
    }
    const bool ok = rows_ok && x >= HW && x < W - HW;
    cen[i] = ok ? (uint64_t(lo) | (uint64_t(hi) << 32)) : 0ull;
  }
  const size_t row0 = size_t(y) * W + x0 + xq;
  if (img == 1) {
    DE_GLOBAL uint64_t* cr = gptr(F.cr);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (x0 + xq + i < W) cr[row0 + i] = cen[i];
    return;
  }
  DE_GLOBAL uint4* lrec = gptr(reinterpret_cast<uint4*>(F.lrec));
  DE_GLOBAL uint32_t* ayt = gptr(F.ayt);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int x = x0 + xq + i;
    if (x >= W) break;
    const int lc = int(byte_at(HH, i + HW));
    const uint32_t ax = x ? srf[abs(lc - int(byte_at(HH, i + HW - 1)))] : srf[0];
    const uint32_t ay = y ? srf[abs(lc - int(byte_at(HH - 1, i + HW)))] : srf[0];
    uint4 rec;
    rec.x = uint32_t(cen[i]);
    rec.y = uint32_t(cen[i] >> 32);
    // kTblRow * L (the score table's row offset per intensity unit) and the
    // filter word: 2a in the high half, the complementary weight 65536 - 2a
    // capped at 65535 in the low half (a = 0 then rounds to the same high half)
    rec.z = kTblRow * uint32_t(lc);
    rec.w = ((2u * ax) << 16) | min(65536u - 2u * ax, 65535u);
    lrec[row0 + i] = rec;
    ayt[size_t(x) * H + y] = 2u * ay;
  }
}

// ---------------------------------------------------------------------------
// Stage 2: graded score + H-fwd + H-bwd of one row (k_row, below), and the
// pieces the prior-masked row pass (k_mrow) shares with it. The masked pass
// stages the left record in its packed form z = kTblRow*L | 2a << 16,
// w = 65536 - 2a (the intensity difference masks z's low half first:
// v_sad_u16 would add the high halves' difference too).

// 2*f' = hi16(z) * hi16(t2prev) + acc, with acc = (65536 - 2a) * c' + 32768
// computed off the dependency chain.
__device__ __forceinline__ uint32_t rf_chain_z(uint32_t z, uint32_t t2prev, uint32_t acc) {
  uint32_t r;
  asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[1,1,0,0]" : "=v"(r) : "v"(z), "v"(t2prev), "v"(acc));
  return r;
}

// popcount(a) + acc in one instruction (kept as written: the compiler would
// otherwise sum two counts and shift in a fused add-shift plus another add)
__device__ __forceinline__ uint32_t bcnt_acc(uint32_t a, uint32_t acc) {
  uint32_t r;
  asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(acc));
  return r;
}

struct RowLds {
  uint32_t hb;               // this wave's table address / 2 (added by the first bcnt)
  const uint4* lr;           // left records of the row (z low half = kTblRow * L)
  const uint64_t* crs;       // right census of the row
  const uint32_t* r2s;       // kTblRow * R of the row (a word each)
  uint32_t Tb;               // kTblRow * T
};

// Score table as staged for the row pass: intensity-difference-major, rows of
// kTblRow bytes (66 u16, the census distance 0..62 in a row). A wave's 64
// lanes share the intensity row when their differences saturate at T (the
// common case), and the census distance then spreads them over the banks; the
// padded stride (33 words) also separates lanes of equal distance on
// different rows. Census-distance-major rows of 2^s entries put every
// saturated lane on one bank of each row.
// Score-table read. Each wave's table starts its LDS region, at address
// 2*hb (k_row declares no static LDS, so its dynamic LDS starts at address 0);
// hb enters as the first popcount's accumulator, so (h + hb) * 2 +
// kTblRow*min(|L - R|, T) is the address itself: one shift-add, no base add.
__device__ __forceinline__ uint32_t row_tbl(uint32_t hh, uint32_t adb) {
  return *reinterpret_cast<const __attribute__((address_space(3))) uint16_t*>(
      uintptr_t((hh << 1) + adb));
}

__device__ __forceinline__ uint32_t row_score(const RowLds& q, const uint4& rec, uint64_t cr,
                                              uint32_t rb) {
  const uint32_t h = bcnt_acc(rec.y ^ uint32_t(cr >> 32), __popc(rec.x ^ uint32_t(cr)) + q.hb);
  const uint32_t adb = min(sad_u32(rec.z & 0xFFFFu, rb, 0u), q.Tb);
  return row_tbl(h, adb);
}

// LDS operands of 8 columns (stage 1) and their scores with the record words
// the chain needs (stage 2; the table reads may still be in flight).
struct RowRaw {
  uint4 rec[8];
  uint64_t cr[8];
  uint32_t r2[8];
};
struct RowBlk {
  uint32_t c[8];
  uint32_t z[8];
  uint32_t w[8];
};

__device__ __forceinline__ void row_score_raw(const RowLds& q, const RowRaw& R, RowBlk& S) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    S.c[u] = row_score(q, R.rec[u], R.cr[u], R.r2[u]);
    S.z[u] = R.rec[u].z;
    S.w[u] = R.rec[u].w;
  }
}

// Row-pass context of one wave: volume row, forward scratch, transpose tile.
struct RowOut {
  DE_GLOBAL uint8_t* vrow;   // volume, column 0 of this row and group (wave-uniform)
  DE_GLOBAL uint8_t* frow;   // forward scratch of this row and group
  uint16_t* tile;            // LDS, two tiles of [8][64] u16
  uint32_t lane;
  uint32_t so;               // lane's store offset within 8 columns: (lane/8)*XB + (lane%8)*16
  int W;
};

// Final values of columns 8b..8b+7 (read from a tile: lane l holds
// disparities 8(l%8)..8(l%8)+7 of column 8b + l/8) to the volume, 16 bytes per
// lane.
template <uint32_t XB>
__device__ __forceinline__ void row_store8(const RowOut& o, int b, const uint4& v) {
  *reinterpret_cast<DE_GLOBAL uint4*>(o.vrow + size_t(8 * b) * XB + o.so) = v;
}

// The same from tile 0, right away; with CHECK, columns at or beyond W are not
// written.
template <uint32_t XB, bool CHECK>
__device__ __forceinline__ void row_flush(const RowOut& o, int b) {
  const uint4 v = *reinterpret_cast<const uint4*>(o.tile + o.lane * 8);
  const int xc = 8 * b + int(o.lane >> 3);
  if (!CHECK || xc < o.W) row_store8<XB>(o, b, v);
}

// Coefficient words (z, w) of columns x0+1 .. x0+8, read before any tile write
// of the block that uses them (the compiler cannot tell the tile from the
// records, so a read issued after a tile write would wait for it).
__device__ __forceinline__ void row_coef8(uint2 (&zw)[8], const uint4* lr, int x0) {
#pragma unroll
  for (int j = 0; j < 8; ++j) zw[j] = *reinterpret_cast<const uint2*>(&lr[x0 + j + 1].z);
}

template <int P>
__device__ __forceinline__ void row_load_blocks(uint4 (&v)[P], const RowOut& o, int b) {
#pragma unroll
  for (int k = 0; k < P; ++k)
    v[k] = *reinterpret_cast<const DE_GLOBAL uint4*>(o.frow + size_t(b - k) * 1024 + o.lane * 16);
}

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
// block n). The backward scan's final values are transposed through two
// alternating 1 KB LDS tiles and leave as 16-byte stores into the [y][x][Dpad]
// volume, lane l writing disparities 8(l%8)..8(l%8)+7 of column 8b + l/8.

// f(integral_constant<int, i>) for i = 0 .. N-1, expanded at compile time
// whatever the unroller's size limits (every register index is a constant).
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

constexpr int kRowVBlocks = 16;      // forward blocks in VGPRs (64 registers)
constexpr int kRowABlocks = 60;      // forward blocks in AGPRs (240 registers)
constexpr int kRowRegBlocks = kRowVBlocks + kRowABlocks;
constexpr int kRowTailSlots = 3;     // LDS slots of the blocks from the end of the aligned range
constexpr int kRowPad = 16;          // LDS record arrays run this far past W

// acc = lo16(w) * lo16(c) + k
__device__ __forceinline__ uint32_t mad_lo16(uint32_t w, uint32_t c, uint32_t k) {
  uint32_t r;
  asm("v_mad_u32_u16 %0, %1, %2, %3" : "=v"(r) : "v"(w), "v"(c), "v"(k));
  return r;
}
// acc = lo16(w) * hi16(f2) + k
__device__ __forceinline__ uint32_t mad_lohi16(uint32_t w, uint32_t f2, uint32_t k) {
  uint32_t r;
  asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[0,1,0,0]" : "=v"(r) : "v"(w), "v"(f2), "v"(k));
  return r;
}
__device__ __forceinline__ uint32_t to_acc(uint32_t v) {
  uint32_t a;
  asm("v_accvgpr_write_b32 %0, %1" : "=a"(a) : "v"(v));
  return a;
}
__device__ __forceinline__ uint32_t from_acc(uint32_t a) {
  uint32_t v;
  asm("v_accvgpr_read_b32 %0, %1" : "=v"(v) : "a"(a));
  return v;
}

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

// H-bwd over P blocks b, b-1, ..., b-P+1 from their packed forward words. The
// next block's coefficients are read before the current block's tile writes;
// the tiles alternate, and a block's 16-byte stores leave after the next
// block's chain.
template <int P, uint32_t XB>
__device__ __forceinline__ uint32_t rr_back_blocks(const uint4 (&v)[P], uint32_t t2,
                                                   const uint4* lr, int b, const RowOut& o,
                                                   uint32_t K) {
  uint32_t wc[8], wn[8];
  uint4 pend = make_uint4(0u, 0u, 0u, 0u);
  rr_coef8(wc, lr, 8 * b);
#pragma unroll
  for (int k = 0; k < P; ++k) {
    if (k + 1 < P) rr_coef8(wn, lr, 8 * (b - k - 1));
    uint16_t* const tile = o.tile + (k & 1) * 8 * kWave;
    uint32_t t[8];
    rr_chain_back(wc, v[k], t2, K, t);
    t2 = t[0];
#pragma unroll
    for (int j = 7; j >= 0; --j) tile[j * kWave + o.lane] = uint16_t(t[j] >> 16);
    if (k > 0) row_store8<XB>(o, b - k + 1, pend);
    pend = *reinterpret_cast<const uint4*>(tile + o.lane * 8);
#pragma unroll
    for (int j = 0; j < 8; ++j) wc[j] = wn[j];
  }
  row_store8<XB>(o, b - P + 1, pend);
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
// then per wave two transpose tiles and nl forward slots of 1 KB.
__host__ __device__ inline size_t row_ws_fixed(int W, int T, int G) {
  const size_t tb = (size_t(T + 1) * kTblRow + 15) & ~size_t(15);
  return tb + 2 * size_t(W + kRowPad) * 16 + size_t(G) * 2048;
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
  uint8_t* const wbase = reinterpret_cast<uint8_t*>(rrec + (W + kRowPad)) + g * (2048 + nl * 1024);
  o.tile = reinterpret_cast<uint16_t*>(wbase);
  uint8_t* const lf = wbase + 2048;                      // this wave's forward slots
  {
    // entry (a, h) = score of census distance h, intensity difference a <= T;
    // the frame's table holds it at (h << sc_shift) + a
    uint16_t* tbl = reinterpret_cast<uint16_t*>(smem);
    const DE_CONST uint16_t* sc = cptr(F.sc);
    for (int i = threadIdx.x; i < ntt; i += 64 * G) {
      const int a = i / int(kTblRow / 2), h = i - a * int(kTblRow / 2);
      tbl[i] = h <= nb ? sc[(h << F.sc_shift) + a] : uint16_t(kOff);
    }
    const DE_CONST uint4* lrg = cptr(reinterpret_cast<const uint4*>(F.lrec)) + size_t(y) * W;
    const DE_CONST uint64_t* crg = cptr(F.cr) + size_t(y) * W;
    const DE_CONST uint8_t* Rg = cptr(F.R) + size_t(y) * W;
    for (int x = threadIdx.x; x < W + kRowPad; x += 64 * G) {
      uint4 rec = make_uint4(0u, 0u, 0u, 0u), rv = make_uint4(0u, 0u, 0u, 0u);
      if (x < W) {
        rec = lrg[x];
        const uint64_t c = crg[x];
        rv = make_uint4(uint32_t(c), uint32_t(c >> 32), kTblRow * Rg[x], 0u);
      }
      lr[x] = rec;
      rrec[x] = rv;
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
  for (; x < xb; ++x) {                                  // ragged band: per-lane window
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
  RrRaw R[2];
  RrBlk S[2];
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
      rr_load_raw(lr, rrec, R[0], 8 * k + 16, d);
      rr_score_raw(R[1], S[1], Tb);
      t2 = rr_chain(S[0], t2, K, v);
      put(k, v);
      rr_load_raw(lr, rrec, R[1], 8 * k + 24, d);
      rr_score_raw(R[0], S[0], Tb);
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
    static_for<NT>([&](auto J) __attribute__((always_inline)) {
      constexpr int j = decltype(J)::value;
      if (j >= j0) {
        const int xj = 8 * (Bb + j);
        if constexpr (j + 2 < NT) rr_load_raw(lr, rrec, R[j & 1], xj + 16, d);
        if constexpr (j + 1 < NT) rr_score_raw(R[(j + 1) & 1], S[(j + 1) & 1], Tb);
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
      }
    });
    x = 8 * B1;
  }
  for (; x < xend; ++x) {
    const uint4 rec = lr[x];
    const uint3 rv = *reinterpret_cast<const uint3*>(&rrec[x - d]);
    t2 = rf_chain_z(rec.w, t2, mad_lo16(rec.w, rr_score(rec, rv, Tb), K));
    store_col(x, t2 >> 16);
  }
  for (; x < W; ++x) {                                   // right border: zero scores
    const uint32_t w = lr[x].w;
    t2 = rf_chain_z(w, t2, mad_lo16(w, uint32_t(kOff), K));
    store_col(x, t2 >> 16);
  }

  // ---- H-bwd: g'[W-1] = f'[W-1]; coefficient of x+1 for column x ----
  auto back1 = [&](int xc) {
    const uint32_t fv = xc >= xa ? load_col(xc) : uint32_t(kOff);
    const uint32_t w = lr[xc + 1].w;
    t2 = rf_chain_z(w, t2, mad_lo16(w, fv, K));
    o.tile[(xc & 7) * kWave + lane] = uint16_t(t2 >> 16);
    if ((xc & 7) == 0) row_flush<XB, true>(o, xc >> 3);
  };
  o.tile[((W - 1) & 7) * kWave + lane] = uint16_t(t2 >> 16);
  if (((W - 1) & 7) == 0) row_flush<XB, true>(o, (W - 1) >> 3);
  for (x = W - 2; x >= 8 * B1; --x) back1(x);            // columns from the aligned range's end
  if (nreg > 0) {
    uint4 vr[NV];                                        // block B1 - 1 - k first
#pragma unroll
    for (int k = 0; k < NV; ++k) vr[k] = fr[NV - 1 - k];
    t2 = rr_back_blocks<NV, XB>(vr, t2, lr, B1 - 1, o, K);
    static_for<NA / 4>([&](auto C) __attribute__((always_inline)) {
      constexpr int c = NA / 4 - 1 - decltype(C)::value;   // chunks from the top down
      if (4 * c >= j0) {
        uint4 v[4];
        static_for<4>([&](auto Q) __attribute__((always_inline)) {
          constexpr int q = decltype(Q)::value, j = 4 * c + 3 - q;
          v[q] = make_uint4(from_acc(ag[4 * j]), from_acc(ag[4 * j + 1]),
                            from_acc(ag[4 * j + 2]), from_acc(ag[4 * j + 3]));
        });
        t2 = rr_back_blocks<4, XB>(v, t2, lr, Bb + 4 * c + 3, o, K);
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
  // blocks below the registers: LDS [max(blo, bL), Bd), then global [blo, bL)
  const int blo = min((xa + 7) >> 3, Bd);
  t2 = rr_back_range<4, XB>(
      [&](int bk) { return *reinterpret_cast<const uint4*>(lds_block(bk)); }, t2, lr, Bd - 1,
      max(blo, bL), o, K);
  t2 = rr_back_range<P, XB>(
      [&](int bk) { return *reinterpret_cast<const DE_GLOBAL uint4*>(glb_block(bk)); }, t2, lr,
      min(bL, Bd) - 1, blo, o, K);
  for (x = 8 * blo - 1; x >= 0; --x) back1(x);
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
    const DE_CONST uint16_t* sc = cptr(F.sc);
    for (int i = lane; i < ntt; i += kWave) {
      const int a = i / int(kTblRow / 2), h = i - a * int(kTblRow / 2);
      tbl[i] = h <= nb ? sc[(h << F.sc_shift) + a] : uint16_t(kOff);
    }
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
  for (int x = xa; x < xb; ++x) {                        // ragged band: per-lane window
    const uint4 rec = lr[x];
    const bool ok = x >= F.hw + d;
    const int j = ok ? x - d : 0;
    const uint32_t c = ok ? row_score(q, rec, crs[j], r2s[j]) : uint32_t(kOff);
    t2 = rf_chain_z(rec.z, t2, __umul24(rec.w, c) + 32768u);
    *fwd_at(x) = uint16_t(t2 >> 16);
  }
  const int xm0 = min((xb + 7) & ~7, xend);              // aligned blocks [xm0, xm0 + 8 nblk)
  const int nblk_all = max(0, ((xend & ~7) - xm0) >> 3);
  // the last kRowSRegBlocks aligned blocks [E - NB, E) go to registers when at
  // least two blocks precede them
  constexpr int NB = kRowSRegBlocks;
  const bool use_reg = nblk_all >= NB + 2;
  const int nblk = use_reg ? nblk_all - NB : nblk_all;   // blocks through the scratch
  const int E = (xm0 >> 3) + nblk_all;
  uint4 fr[NB];
  int x = xb;
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
  const int blo = min(bhi, (xa + 7) >> 3);
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
  for (x = 8 * blo - 1; x >= 0; --x) back1(x);
}

// ---------------------------------------------------------------------------
// Stage 3: V-fwd + V-bwd + exact top-2 with sub-pixel neighbours.
//
// A workgroup owns one column x: G waves, wave g holding disparity group g
// (one chain per lane). V-bwd runs in blocks of 64 rows. While a block is
// filtered, every lane drops its final value into an LDS block (row-major,
// groups 2w and 2w+1 packed in one word, row stride padded by one word so the
// transposed reads hit distinct banks). After a barrier the block is read
// transposed: each row is owned by S lanes (S = 1, 2, 2, 4 for G = 1..4, so
// that every lane of the phase has work), each running the top-2 over its
// share of the row's disparities on packed keys ((v'+1) << 8 | 255-k); the
// shares merge through two lane shuffles per round. Keys order by value then
// smaller k, exactly as the reference's ascending-k strictly-greater
// insertion does, so the merge order does not matter. The neighbours k*-1 and
// k*+1 are two more LDS reads. The V-bwd values never go back to memory
// unless the volume is kept; the output is one 16-byte record per pixel.
//
// Volume accesses are buffer instructions: a wave-uniform resource based at
// the first row of a chunk, the lane's byte offset, and the row step u * (row
// bytes) as a scalar offset the compiler keeps in SGPRs, so a row step costs
// one memory instruction and no address arithmetic. The volume carries
// kVolPadRows rows of padding above and below the frame, so the prefetch runs
// up to 2U rows past either end without bounds checks (the padding is read,
// never used).

__device__ __forceinline__ __amdgpu_buffer_rsrc_t col_rsrc(DE_GLOBAL uint8_t* p) {
  // raw buffer (stride 0); word 3 is the gfx9 data-format field of a raw
  // dword buffer; no range check is relied on (every access is in the volume)
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(p), 0, 0x7FFFFFFF,
                                           0x00020000);
}

// top-2 of the union of two top-2 lists of packed keys
__device__ __forceinline__ void top2_merge(uint32_t& p0, uint32_t& p1, uint32_t q0, uint32_t q1) {
  const uint32_t n1 = max(min(p0, q0), max(p1, q1));
  p0 = max(p0, q0);
  p1 = n1;
}

// NRB: V-fwd blocks of 64 rows kept in registers. Rows yreg .. H-2 (yreg =
// H-1-64*NRB, at least 1 for every frame of the launch) are filtered forward
// into registers instead of the volume, and the first NRB V-bwd blocks read
// them there: 2 * 64 * NRB bytes per lane that neither go to memory nor come
// back. Row H-2-j sits in half j & 1 of register word j >> 1; the code over
// those rows is unrolled, so every slot is a fixed register.
template <int G, int U, bool STORE, int NRB>
__global__ __launch_bounds__(64 * G) void k_col(BatchView b, int base) {
  constexpr int NWL = (G + 1) / 2;           // LDS words per lane per row
  constexpr int RS = kWave * NWL + 1;        // padded row stride (words)
  constexpr int RPW = (kWave + G - 1) / G;   // rows per wave in the transposed phase
  constexpr int S = G == 1 ? 1 : (G == 4 ? 4 : 2);   // lanes per row
  constexpr int NW = kWave * NWL;            // words per row
  constexpr int WPS = NW / S;                // words per lane
  constexpr int NRR = kWave * NRB;           // rows kept in registers
  static_assert(RPW * S <= kWave && NW % S == 0, "transposed phase shape");
  static_assert(kWave % (2 * U) == 0 && 2 * U <= kVolPadRows, "prefetch depth");
  __shared__ uint32_t rb[kWave * RS];
  const int item = blockIdx.x + base;
  const int f = find_frame(b, kStageCol, item);
  const FrameDesc F = b.frames[f];
  const int x = local_item(b, kStageCol, item, f);
  const int W = F.W, H = F.H, Dp = F.Dpad;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t vo = 2u * lane;                         // lane's byte offset in a row
  const int ysb = W * Dp * 2;                            // bytes per volume row
  // wave-uniform address of row 0 of this column and group
  DE_GLOBAL uint8_t* const colb =
      reinterpret_cast<DE_GLOBAL uint8_t*>(gptr(F.vol) + (size_t(x) * Dp + g * kWave));
  auto rs = [&](int y0) { return col_rsrc(colb + ptrdiff_t(y0) * ysb); };
  const DE_CONST uint32_t* ay = cptr(F.ayt) + size_t(x) * H;
  uint16_t* const rb16 = reinterpret_cast<uint16_t*>(rb);
  const int slot = 2 * ((g >> 1) * kWave + lane) + (g & 1);   // u16 index within a row
  uint32_t t = uint32_t(__builtin_amdgcn_raw_buffer_load_b16(rs(0), vo, 0, 0)) << 16;
  const int yreg = H - 1 - NRR;                          // first register row
  uint32_t rr[NRB > 0 ? NRR / 2 : 1];

  // ---- V-fwd over rows 1 .. yend-1 in the volume, U rows in flight ----
  auto ldf = [&](uint32_t (&v)[U], int y0) {             // rows y0 .. y0+U-1
    const __amdgpu_buffer_rsrc_t rr0 = rs(y0);
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = __builtin_amdgcn_raw_buffer_load_b16(rr0, vo, u * ysb, 0);
  };
  {
    auto fwd = [&](const uint32_t (&v)[U], int y0, int n) {
      const __amdgpu_buffer_rsrc_t rr0 = rs(y0);
      uint32_t ab[U];                                    // coefficients of the chunk, one
      __builtin_memcpy(ab, ay + y0, sizeof(ab));         // wide scalar load
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (u < n) {
          const uint32_t a2 = ab[u];
          t = rf_chain(a2, t, rf_acc(a2, v[u]));
          __builtin_amdgcn_raw_buffer_store_b16(uint16_t(t >> 16), rr0, vo, u * ysb, 0);
        }
      }
    };
    const int yend = NRB > 0 ? yreg : H;
    uint32_t A[U], B[U];
    int y = 1;
    ldf(A, y);
    for (; y + 2 * U <= yend; y += 2 * U) {
      ldf(B, y + U);
      fwd(A, y, U);
      ldf(A, y + 2 * U);
      fwd(B, y + U, U);
    }
    if (y + U <= yend) {
      ldf(B, y + U);
      fwd(A, y, U);
      fwd(B, y + U, yend - y - U);
    } else {
      fwd(A, y, yend - y);
    }
  }
  if constexpr (NRB > 0) {
    // ---- V-fwd over the register rows yreg .. H-2, then row H-1 ----
    uint32_t V[2][U];
    ldf(V[0], yreg);
#pragma unroll
    for (int c = 0; c < NRR / U; ++c) {
      const int y0 = yreg + c * U;
      ldf(V[(c + 1) & 1], c + 1 < NRR / U ? y0 + U : H - 1);   // the last brings row H-1
      uint32_t ab[U];
      __builtin_memcpy(ab, ay + y0, sizeof(ab));
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t a2 = ab[u];
        t = rf_chain(a2, t, rf_acc(a2, V[c & 1][u]));
        const int j = NRR - 1 - (c * U + u);             // row y0 + u = H-2-j; odd j first
        if (j & 1) rr[j >> 1] = t & 0xFFFF0000u;
        else rr[j >> 1] |= t >> 16;
      }
    }
    const uint32_t a2 = ay[H - 1];
    t = rf_chain(a2, t, rf_acc(a2, V[(NRR / U) & 1][0]));
    if (STORE) __builtin_amdgcn_raw_buffer_store_b16(uint16_t(t >> 16), rs(H - 1), vo, 0, 0);
  }

  // ---- V-bwd in blocks of 64 rows, transposed top-2 ----
  const int kmax = (x >= F.hw + F.dmin && x < W - F.hw) ? min(F.D - 1, x - F.hw - F.dmin) : -1;
  const int ylo = F.hh, yhi = H - F.hh;                  // rows with candidates
  DE_GLOBAL uint4* rec = gptr(reinterpret_cast<uint4*>(F.cand)) + x;   // row y at rec[y*W]
  if (threadIdx.x == 0) rec[size_t(H - 1) * W] = make_uint4(0u, 0u, 0xFFFFFFFFu, 0u);
  // rows ytop, ytop-1, ..., ytop-U+1 (resource based at the lowest row index)
  auto ldb = [&](uint32_t (&v)[U], int ytop) {
    const __amdgpu_buffer_rsrc_t rr0 = rs(ytop - U + 1);
#pragma unroll
    for (int u = 0; u < U; ++u)
      v[u] = __builtin_amdgcn_raw_buffer_load_b16(rr0, vo, (U - 1 - u) * ysb, 0);
  };
  // one V-bwd row: value v of row ytop - u, coefficient a2 of the row above it
  auto bwd1 = [&](uint32_t v, uint32_t a2, const __amdgpu_buffer_rsrc_t& rr0, int u, int ru) {
    t = rf_chain(a2, t, rf_acc(a2, v));
    const uint32_t vv = t >> 16;
    if (STORE) __builtin_amdgcn_raw_buffer_store_b16(uint16_t(vv), rr0, vo, (U - 1 - u) * ysb, 0);
    rb16[ru * (2 * RS) + slot] = uint16_t(vv);
  };
  auto bwd = [&](const uint32_t (&v)[U], int ytop, int u0, int n) {
    const __amdgpu_buffer_rsrc_t rr0 = rs(ytop - U + 1);
    // coefficient of row ytop-u+1 = ab[U-1-u]: one wide scalar load per chunk
    uint32_t ab[U];
    __builtin_memcpy(ab, ay + (ytop - U + 2), sizeof(ab));
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (u < n) bwd1(v[u], ab[U - 1 - u], rr0, u, u0 + u);
  };
  // transposed-phase ownership: lane = sp * RPW + ri owns share sp of row g*RPW + ri
  const int ri = int(lane) % RPW, sp = int(lane) / RPW;
  const int r = g * RPW + ri;
  const bool kfull = kmax >= NWL * 128 - 1;              // every word of the row valid
  const int nw = min(NW, (max(kmax, 0) >> 7) * kWave + kWave);
  // top-2 of the block rows yb, yb-1, ..., yb-nrow+1 held in rb
  auto top2_block = [&](int yb, int nrow) {
    __syncthreads();
    const int yy = yb - r;
    const bool own = sp < S && r < nrow;
    const bool cand = own && yy >= ylo && yy < yhi && kmax >= 0;
    uint32_t p0 = 0u, p1 = 0u;
    if (cand) {
      const uint32_t* row = rb + r * RS + sp * WPS;
      if (kfull) {
#pragma unroll 8
        for (int jj = 0; jj < WPS; ++jj) {
          const int j = sp * WPS + jj;
          const uint32_t wv = row[jj];
          const int klo = ((j >> 6) << 7) + (j & 63);   // word j holds k = klo and klo + 64
          const uint32_t a = (wv << 8 & 0xFFFF00u) + uint32_t(256 + 255 - klo);
          const uint32_t c = (wv >> 8 & 0xFFFF00u) + uint32_t(256 + 255 - 64 - klo);
          p1 = med3(p0, p1, a);
          p0 = max(p0, a);
          p1 = med3(p0, p1, c);
          p0 = max(p0, c);
        }
      } else {
        // left-border columns (kmax below the group range): bounds per key
#pragma unroll 2
        for (int jj = 0; jj < WPS; ++jj) {
          const int j = sp * WPS + jj;
          if (j >= nw) break;
          const uint32_t wv = row[jj];
          const int klo = ((j >> 6) << 7) + (j & 63);
          const uint32_t key_lo = (wv << 8 & 0xFFFF00u) + uint32_t(256 + 255 - klo);
          const uint32_t key_hi = (wv >> 8 & 0xFFFF00u) + uint32_t(256 + 255 - 64 - klo);
          const uint32_t a = klo <= kmax ? key_lo : 0u;
          const uint32_t c = klo + 64 <= kmax ? key_hi : 0u;
          p1 = med3(p0, p1, a);
          p0 = max(p0, a);
          p1 = med3(p0, p1, c);
          p0 = max(p0, c);
        }
      }
    }
    // merge the S shares of each row (every lane takes part in the shuffles)
#pragma unroll
    for (int m = 1; m < S; m <<= 1) {
      const int src = ((sp ^ m) < S ? (sp ^ m) : sp) * RPW + ri;
      const uint32_t q0 = __shfl(p0, src), q1 = __shfl(p1, src);
      top2_merge(p0, p1, q0, q1);
    }
    if (own && sp == 0) {
      uint32_t nbw = 0xFFFFFFFFu;
      if (cand && p0) {
        const uint32_t* row = rb + r * RS;
        const int ks = key_k(p0);
        auto val = [&](int k) -> uint32_t {
          const uint32_t wv = row[((k >> 7) << 6) + (k & 63)];
          return ((k >> 6) & 1) ? (wv >> 16) : (wv & 0xFFFFu);
        };
        const uint32_t nm = ks >= 1 ? val(ks - 1) : 0xFFFFu;
        const uint32_t np = ks + 1 <= kmax ? val(ks + 1) : 0xFFFFu;
        nbw = nm | (np << 16);
      }
      rec[size_t(yy) * W] = make_uint4(cand ? p0 : 0u, cand ? p1 : 0u, nbw, 0u);
    }
    __syncthreads();
  };
  if constexpr (NRB > 0) {
    // the register blocks: block k holds rows H-2-64k .. H-65-64k, slots 64k ..
#pragma unroll
    for (int k = 0; k < NRB; ++k) {
      const int yb = H - 2 - kWave * k;
#pragma unroll
      for (int q = 0; q < kWave / U; ++q) {
        const int ytop = yb - q * U;
        const __amdgpu_buffer_rsrc_t rr0 = rs(ytop - U + 1);
        uint32_t ab[U];
        __builtin_memcpy(ab, ay + (ytop - U + 2), sizeof(ab));
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int j = kWave * k + q * U + u;
          const uint32_t v = (j & 1) ? (rr[j >> 1] >> 16) : (rr[j >> 1] & 0xFFFFu);
          bwd1(v, ab[U - 1 - u], rr0, u, q * U + u);
        }
      }
      top2_block(yb, kWave);
    }
  }
  const int yb0 = H - 2 - NRR;                           // first block from the volume
  uint32_t A[U], B[U];
  ldb(A, yb0);
  for (int yb = yb0; yb >= 0; yb -= kWave) {            // block: rows yb, yb-1, ..., yb-63
    const int nrow = min(kWave, yb + 1);
    int u0 = 0;
    for (; u0 + 2 * U <= nrow; u0 += 2 * U) {
      ldb(B, yb - u0 - U);
      bwd(A, yb - u0, u0, U);
      ldb(A, yb - u0 - 2 * U);
      bwd(B, yb - u0 - U, u0 + U, U);
    }
    if (u0 < nrow) {                                     // last block: fewer than 2U rows left
      if (u0 + U <= nrow) {
        ldb(B, yb - u0 - U);
        bwd(A, yb - u0, u0, U);
        bwd(B, yb - u0 - U, u0 + U, nrow - u0 - U);
      } else {
        bwd(A, yb - u0, u0, nrow - u0);
      }
    }
    top2_block(yb, nrow);
  }
}

// ---------------------------------------------------------------------------
// Standalone census of one image into a u64 plane (the census_transform API).
template <int HW, int HH>
__global__ __launch_bounds__(256) void k_census_plane(const uint8_t* __restrict__ img, int W,
                                                      int H, uint64_t* __restrict__ out) {
  constexpr int TW = kPrepTileW + 2 * HW;
  constexpr int TH = kPrepTileH + 2 * HH;
  __shared__ uint8_t s[TH][TW];
  const int ntx = (W + kPrepTileW - 1) / kPrepTileW;
  const int x0 = (blockIdx.x % ntx) * kPrepTileW, y0 = (blockIdx.x / ntx) * kPrepTileH;
  for (int i = threadIdx.x; i < TW * TH; i += 256) {
    const int ty = i / TW, tx = i - ty * TW;
    const int gx = x0 - HW + tx, gy = y0 - HH + ty;
    s[ty][tx] = (gx >= 0 && gx < W && gy >= 0 && gy < H) ? img[size_t(gy) * W + gx] : 0;
  }
  __syncthreads();
  const int lx = threadIdx.x & 63;
  for (int r = threadIdx.x >> 6; r < kPrepTileH; r += 4) {
    const int x = x0 + lx, y = y0 + r;
    if (x >= W || y >= H) continue;
    uint64_t v = 0;
    if (x >= HW && x < W - HW && y >= HH && y < H - HH) {
      v = census_at<HW, HH, TW>(s, r + HH, lx + HW);
    }
    out[size_t(y) * W + x] = v;
  }
}

// ---------------------------------------------------------------------------
// Host launchers (called from the C-ABI translation unit).
size_t row_lds_bytes(int W, int nbits, int sc_shift) {
  const int ntbl = (nbits + 1) << sc_shift;
  // table (both row passes stage it transposed: at most 2^sc_shift rows of
  // 132 bytes, or the frame's table size if larger) | two row-pass transpose
  // tiles | records, right census, intensity
  const size_t tb = std::max(size_t(ntbl) * 2, size_t(132) << sc_shift);
  // (the masked row pass pads each row array by 8 entries)
  return ((tb + 15) & ~size_t(15)) + 2048 + size_t(W + 8) * (16 + 8 + 4);
}

hipError_t launch_prep(const BatchView& b, int items, int hw, int hh, hipStream_t s) {
  if (hw == 3 && hh == 3) hipLaunchKernelGGL((k_prep<3, 3>), dim3(items), dim3(256), 0, s, b);
  else if (hw == 4 && hh == 3) hipLaunchKernelGGL((k_prep<4, 3>), dim3(items), dim3(256), 0, s, b);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

// Prefetch-depth variants (chosen per shape bucket by de_autotune).
// row pass: resident (global-scratch prefetch 4, 16 blocks), streaming (8, 16)
const int kRowVariants[kNumRowVariants] = {4, 16, 8, 16};
const int kColVariants[kNumColVariants] = {8, 16, 32};

// Device LDS per workgroup and CU count (queried once).
static void device_shape(int* lds, int* cus) {
  static int l = 0, c = 0;
  static std::once_flag once;
  std::call_once(once, [] {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&l, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
      l = 64 * 1024;
      c = 1;
    }
  });
  *lds = l;
  *cus = c;
}

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

// Variants: 0, 1 the register-resident pass (global-scratch prefetch 4 or 16
// blocks), 2, 3 the streaming pass (backward prefetch 8 or 16 blocks).
template <int G>
static hipError_t launch_row_g(const BatchView& b, int base, int items, int maxW, int T,
                               size_t lds_wave, int variant, hipStream_t s) {
  switch (variant) {
    case 0: return launch_row_gp<G, 4>(b, base, items, maxW, T, true, s);
    case 1: return launch_row_gp<G, 16>(b, base, items, maxW, T, true, s);
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

template <int G, bool STORE, int NRB>
static void launch_col_gn(const BatchView& b, int base, int items, int variant, hipStream_t s) {
  switch (variant) {
    case 0: hipLaunchKernelGGL((k_col<G, 8, STORE, NRB>), dim3(items), dim3(64 * G), 0, s, b, base); break;
    case 1: hipLaunchKernelGGL((k_col<G, 16, STORE, NRB>), dim3(items), dim3(64 * G), 0, s, b, base); break;
    default: hipLaunchKernelGGL((k_col<G, 32, STORE, NRB>), dim3(items), dim3(64 * G), 0, s, b, base); break;
  }
}

// Register blocks by the frames' height and the launch's size. A launch of at
// most two waves per SIMD keeps 3 blocks (192 rows, ~130 more VGPRs) when every
// frame has H >= 194; a larger one 2 blocks (128 rows), which leave room for
// more waves per SIMD; frames under 130 rows none. Measured at 848x480x128
// (1696 waves): column pass 66 us without, 62 with 2 blocks, 60 with 3; four
// frames in flight 0.141 / 0.137 ms a frame with 2 / 3 blocks. The 64-pair
// batch ran 10.9 ms with 2 blocks, 11.9 with 3. Four blocks (256 VGPRs, AGPR
// spills) took 76 us.
template <int G, bool STORE>
static void launch_col_g(const BatchView& b, int base, int items, int variant, int min_h,
                         hipStream_t s) {
  int lds = 0, cus = 1;
  device_shape(&lds, &cus);
  const bool small = size_t(items) * G <= size_t(8) * cus;
  if (small && min_h >= 3 * kWave + 2) launch_col_gn<G, STORE, 3>(b, base, items, variant, s);
  else if (min_h >= 2 * kWave + 2) launch_col_gn<G, STORE, 2>(b, base, items, variant, s);
  else launch_col_gn<G, STORE, 0>(b, base, items, variant, s);
}

// min_h: the smallest H of the launch's frames (sets the register blocks).
hipError_t launch_col(const BatchView& b, int base, int items, int G, bool store, int variant,
                      int min_h, hipStream_t s) {
  if (variant < 0 || variant >= kNumColVariants) return hipErrorInvalidValue;
  switch (G * 2 + (store ? 1 : 0)) {
    case 2: launch_col_g<1, false>(b, base, items, variant, min_h, s); break;
    case 3: launch_col_g<1, true>(b, base, items, variant, min_h, s); break;
    case 4: launch_col_g<2, false>(b, base, items, variant, min_h, s); break;
    case 5: launch_col_g<2, true>(b, base, items, variant, min_h, s); break;
    case 6: launch_col_g<3, false>(b, base, items, variant, min_h, s); break;
    case 7: launch_col_g<3, true>(b, base, items, variant, min_h, s); break;
    case 8: launch_col_g<4, false>(b, base, items, variant, min_h, s); break;
    case 9: launch_col_g<4, true>(b, base, items, variant, min_h, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_census_plane(const uint8_t* img, int W, int H, int hw, int hh,
                               uint64_t* out, hipStream_t s) {
  const int tiles = ((W + kPrepTileW - 1) / kPrepTileW) * ((H + kPrepTileH - 1) / kPrepTileH);
  if (hw == 3 && hh == 3)
    hipLaunchKernelGGL((k_census_plane<3, 3>), dim3(tiles), dim3(256), 0, s, img, W, H, out);
  else if (hw == 4 && hh == 3)
    hipLaunchKernelGGL((k_census_plane<4, 3>), dim3(tiles), dim3(256), 0, s, img, W, H, out);
  else if (hw == 2 && hh == 2)
    hipLaunchKernelGGL((k_census_plane<2, 2>), dim3(tiles), dim3(256), 0, s, img, W, H, out);
  else if (hw == 1 && hh == 1)
    hipLaunchKernelGGL((k_census_plane<1, 1>), dim3(tiles), dim3(256), 0, s, img, W, H, out);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Exact reverse match (--lrc, core/tools/de_dense.cpp:1438-1453, 1581-1597):
// for every right pixel, the disparity index whose final filtered score is the
// largest over all left pixels claiming it (x = xr + d), the smallest index
// among equals; then the residual plane |d(x) - (dmin + that index at
// x - round(d(x)))|. Reads the volume the column pass stored back.
__global__ __launch_bounds__(256) void k_lrc_exact(BatchView b) {
  const int item = blockIdx.x;
  const int f = find_frame(b, kStageSolve, item);
  const FrameDesc F = b.frames[f];
  if (!F.lrc) return;
  const int y = local_item(b, kStageSolve, item, f);
  const int W = F.W, Dp = F.Dpad, D = F.D;
  const int lane = threadIdx.x & 63;
  const bool yok = y >= F.hh && y < F.H - F.hh;
  for (int xr = threadIdx.x >> 6; xr < W; xr += 4) {
    uint32_t key = 0u;
    if (yok && xr >= F.hw) {
      for (int k = lane; k < D; k += kWave) {
        const int x = xr + F.dmin + k;
        if (x >= W - F.hw) break;
        const uint32_t v = F.vol[(size_t(y) * W + x) * Dp + k];
        key = max(key, pack_key(v, k));
      }
    }
    for (int o = 1; o < kWave; o <<= 1) key = max(key, uint32_t(__shfl_xor(int(key), o)));
    if (lane == 0) F.rdisp[size_t(y) * W + xr] = key ? int16_t(key_k(key)) : int16_t(-1);
  }
}

__global__ __launch_bounds__(256) void k_lrc_plane(BatchView b) {
  const int item = blockIdx.x;
  const int f = find_frame(b, kStageSolve, item);
  const FrameDesc F = b.frames[f];
  if (!F.lrc || !F.cand6) return;
  const int y = local_item(b, kStageSolve, item, f);
  const int W = F.W;
  const size_t WH = size_t(W) * F.H;
  for (int x = threadIdx.x; x < W; x += 256) {
    const size_t i = size_t(y) * W + x;
    const float dl = F.disp[i];
    float r = __int_as_float(0x7fc00000);
    if (isfinite(dl)) {
      const int xr = x - int(lroundf(dl));
      if (xr >= 0 && xr < W) {
        const int rd = F.rdisp[size_t(y) * W + xr];
        if (rd >= 0) r = fabsf(dl - float(F.dmin + rd));
      }
    }
    F.cand6[4 * WH + i] = r;
  }
}

hipError_t launch_lrc_exact(const BatchView& b, int items, hipStream_t s) {
  hipLaunchKernelGGL(k_lrc_exact, dim3(items), dim3(256), 0, s, b);
  return hipGetLastError();
}

hipError_t launch_lrc_plane(const BatchView& b, int items, hipStream_t s) {
  hipLaunchKernelGGL(k_lrc_plane, dim3(items), dim3(256), 0, s, b);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Prior-masked sweep (--prior / --c2f, core/tools/de_dense.cpp:797-853,
// 1156-1296). Plane k is evaluated only on its rectangle: the bounding box of
// the pixels whose band |d - round(prior)| <= band admits it, padded by 8 and
// clipped to the score region. All four filter passes are restricted to that
// rectangle and start from its edge values (the reference's rect filter), and
// a score enters the candidates only inside its row's interval. Lane = plane,
// as in the full sweep; lanes whose plane does not cover the row (column)
// idle through it.

// Rounded prior per pixel, clamped so that dc +- band cannot overflow.
__global__ __launch_bounds__(256) void k_mask_dc(const float* __restrict__ prior, int n,
                                                 int* __restrict__ dc) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float r = roundf(prior[i]);   // std::lround: halves away from zero
  dc[i] = r >= 1e9f ? 1000000000 : (r <= -1e9f ? -1000000000 : int(r));
}

// One workgroup per row: the interval [mx0, mx1] of the row's pixels whose
// band admits each plane (mx0 = W, mx1 = -1 when none), by LDS min/max.
__global__ __launch_bounds__(256) void k_mask_iv(const FrameDesc* __restrict__ fp,
                                                 const int* __restrict__ dc, int band,
                                                 int* __restrict__ mx0, int* __restrict__ mx1) {
  const FrameDesc F = *fp;
  __shared__ int lo[kMaxD], hi[kMaxD];
  const int y = blockIdx.x;
  for (int k = threadIdx.x; k < F.D; k += 256) {
    lo[k] = F.W;
    hi[k] = -1;
  }
  __syncthreads();
  const int* row = dc + size_t(y) * F.W;
  for (int x = threadIdx.x; x < F.W; x += 256) {
    const long long c = row[x];
    const long long k0 = max(c - band, (long long)F.dmin) - F.dmin;
    const long long k1 = min(c + band, (long long)(F.dmin + F.D - 1)) - F.dmin;
    for (long long k = k0; k <= k1; ++k) {
      atomicMin(&lo[k], x);
      atomicMax(&hi[k], x);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < F.D; k += 256) {
    mx0[size_t(y) * F.D + k] = lo[k];
    mx1[size_t(y) * F.D + k] = hi[k];
  }
}

// One workgroup per plane: the bounding box of its intervals and the filtered
// rectangle (y0, y1, x0, x1 half-open; y1 = y0 = 0 when the plane is skipped).
__global__ __launch_bounds__(256) void k_mask_rect(const FrameDesc* __restrict__ fp,
                                                   const int* __restrict__ mx0,
                                                   const int* __restrict__ mx1,
                                                   int4* __restrict__ box, int4* __restrict__ rect) {
  const FrameDesc F = *fp;
  const int k = blockIdx.x;
  int ylo = F.H, yhi = -1, xlo = F.W, xhi = -1;
  for (int y = threadIdx.x; y < F.H; y += 256) {
    const int a = mx0[size_t(y) * F.D + k], b = mx1[size_t(y) * F.D + k];
    if (b < a) continue;
    ylo = min(ylo, y);
    yhi = max(yhi, y);
    xlo = min(xlo, a);
    xhi = max(xhi, b);
  }
  __shared__ int red[4][256 / kWave];
  for (int o = 1; o < kWave; o <<= 1) {
    ylo = min(ylo, __shfl_xor(ylo, o));
    yhi = max(yhi, __shfl_xor(yhi, o));
    xlo = min(xlo, __shfl_xor(xlo, o));
    xhi = max(xhi, __shfl_xor(xhi, o));
  }
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[0][w] = ylo;
    red[1][w] = yhi;
    red[2][w] = xlo;
    red[3][w] = xhi;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int q = 1; q < 256 / kWave; ++q) {
    ylo = min(ylo, red[0][q]);
    yhi = max(yhi, red[1][q]);
    xlo = min(xlo, red[2][q]);
    xhi = max(xhi, red[3][q]);
  }
  constexpr int kPad = 8;   // MPAD: real scores this far outside the intervals
  const int d = F.dmin + k;
  const int y0 = max(F.hh, ylo - kPad), y1 = min(F.H - F.hh, yhi + 1 + kPad);
  const int x0 = max(F.hw + d, xlo - kPad), x1 = min(F.W - F.hw, xhi + 1 + kPad);
  const bool skip = yhi < ylo || y1 - y0 < 1 || x1 - x0 < 2;
  box[k] = make_int4(ylo, yhi, xlo, xhi);
  rect[k] = skip ? make_int4(0, 0, 0, 0) : make_int4(y0, y1, x0, x1);
}

// (the results are wave-uniform; readfirstlane tells the compiler, so loads
// addressed by them are scalar)
__device__ __forceinline__ int wave_min(int v) {
  for (int o = 1; o < kWave; o <<= 1) v = min(v, __shfl_xor(v, o));
  return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_max(int v) {
  for (int o = 1; o < kWave; o <<= 1) v = max(v, __shfl_xor(v, o));
  return __builtin_amdgcn_readfirstlane(v);
}

// Score + H-fwd + H-bwd of one row and 64 planes, each plane on its own
// rectangle's x range [xs, xe), built on the full row pass's machinery
// (blocked forward scratch, three-stage forward pipeline, 16-byte transfers,
// transposed tiles). The wave runs the chain over the whole span of its
// planes' ranges in aligned 8-column blocks; a lane restarts it at its own
// first column (f' = score there) and, scanning back, at its last (g' = f').
// Values outside a plane's rectangle are computed and written but never read
// (the column pass and the insert read inside the rectangles only). The LDS
// row arrays carry 8 entries of padding past W so a block may run past the
// row's end; volume stores of columns >= W are skipped.
__device__ __forceinline__ void mrow_load_raw(const RowLds& q, RowRaw& R, int x0, int d) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int j = max(x0 + u - d, 0);
    R.rec[u] = q.lr[x0 + u];
    R.cr[u] = q.crs[j];
    R.r2[u] = q.r2s[j];
  }
}

__device__ __forceinline__ uint32_t mrow_chain_fwd(const RowBlk& S, uint32_t t2, int x0, int xs,
                                                   DE_GLOBAL uint8_t* fw) {
  uint32_t t[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const uint32_t tn = rf_chain_z(S.z[u], t2, __umul24(S.w[u], S.c[u]) + 32768u);
    t2 = x0 + u == xs ? S.c[u] << 16 : tn;
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

// H-bwd over P blocks b, b-1, ..., b-P+1 (as rr_back_blocks), restarting at
// column xl = xe - 1 of each lane; volume stores of columns >= W are skipped.
template <int P, uint32_t XB>
__device__ __forceinline__ uint32_t mrow_back_blocks(const uint4 (&v)[P], uint32_t t2,
                                                     const uint4* lr, int b, int xl,
                                                     const RowOut& o) {
  uint2 zc[8], zn[8];
  uint4 pend = make_uint4(0u, 0u, 0u, 0u);
  row_coef8(zc, lr, 8 * b);
  auto store = [&](int bb, const uint4& val) {
    if (8 * bb + int(o.lane >> 3) < o.W) row_store8<XB>(o, bb, val);
  };
#pragma unroll
  for (int k = 0; k < P; ++k) {
    if (k + 1 < P) row_coef8(zn, lr, 8 * (b - k - 1));
    uint16_t* const tile = o.tile + (k & 1) * 8 * kWave;
    const uint32_t wv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
    const int x0 = 8 * (b - k);
#pragma unroll
    for (int j = 7; j >= 0; --j) {
      const uint32_t f = (j & 1) ? (wv[j >> 1] >> 16) : (wv[j >> 1] & 0xFFFFu);
      const uint32_t tn = rf_chain_z(zc[j].x, t2, __umul24(zc[j].y, f) + 32768u);
      t2 = x0 + j == xl ? f << 16 : tn;
      tile[j * kWave + o.lane] = uint16_t(t2 >> 16);
    }
    if (k > 0) store(b - k + 1, pend);
    pend = *reinterpret_cast<const uint4*>(tile + o.lane * 8);
#pragma unroll
    for (int j = 0; j < 8; ++j) zc[j] = zn[j];
  }
  store(b - P + 1, pend);
  return t2;
}

template <int G>
__global__ __launch_bounds__(64) void k_mrow(const FrameDesc* __restrict__ fp,
                                             const int4* __restrict__ rect) {
  constexpr int Dp = G * kWave;
  constexpr uint32_t XB = 2u * Dp;                       // bytes per volume column
  constexpr int P = 8;                                   // H-bwd blocks in flight
  extern __shared__ __align__(16) uint8_t smem[];
  const FrameDesc F = *fp;
  const int W = F.W;
  const int y = blockIdx.x / G, g = blockIdx.x - y * G;
  const uint32_t lane = threadIdx.x & 63;
  const int k = g * kWave + int(lane);
  const int d = F.dmin + k;
  const int4 r = k < F.D ? rect[k] : make_int4(0, 0, 0, 0);
  const bool act = k < F.D && y >= r.x && y < r.y;
  const int xs = act ? r.z : W, xe = act ? r.w : 0;
  const int xmin = wave_min(xs), xmax = wave_max(xe);
  if (xmin >= xmax) return;   // uniform over the wave
  RowOut o;
  o.vrow = reinterpret_cast<DE_GLOBAL uint8_t*>(gptr(F.vol) + (size_t(y) * W * Dp + g * kWave));
  o.lane = lane;
  o.so = (lane >> 3) * XB + (lane & 7) * 16;
  o.W = W;
  // LDS as the full row pass stages it, with 8 entries of padding per row array
  const int nb = (2 * F.hw + 1) * (2 * F.hh + 1) - 1;
  const int T = F.ad_T;
  const int ntt = (T + 1) * int(kTblRow / 2);
  const int Wp = W + 8;
  o.tile = reinterpret_cast<uint16_t*>(smem + ((ntt * 2 + 15) & ~15));
  uint4* lr = reinterpret_cast<uint4*>(o.tile + 2 * 8 * kWave);
  uint64_t* crs = reinterpret_cast<uint64_t*>(lr + Wp);
  uint32_t* r2s = reinterpret_cast<uint32_t*>(crs + Wp);
  {
    uint16_t* tbl = reinterpret_cast<uint16_t*>(smem);
    for (int i = lane; i < ntt; i += kWave) {
      const int a = i / int(kTblRow / 2), h = i - a * int(kTblRow / 2);
      tbl[i] = h <= nb ? F.sc[(h << F.sc_shift) + a] : uint16_t(kOff);
    }
    const uint4* lrg = reinterpret_cast<const uint4*>(F.lrec) + size_t(y) * W;
    const uint64_t* crg = F.cr + size_t(y) * W;
    const uint8_t* Rg = F.R + size_t(y) * W;
    for (int x = lane; x < Wp; x += kWave) {
      uint4 rec = make_uint4(0u, 0u, 0u, 65536u);
      uint64_t cv = 0;
      uint32_t rv = 0;
      if (x < W) {
        rec = lrg[x];
        // the masked pass keeps the packed form: kTblRow * L | 2a << 16, 65536 - 2a
        rec.z = rec.z | (rec.w & 0xFFFF0000u);
        rec.w = 65536u - (rec.w >> 16);
        cv = crg[x];
        rv = kTblRow * Rg[x];
      }
      lr[x] = rec;
      crs[x] = cv;
      r2s[x] = rv;
    }
  }
  __syncthreads();
  RowLds q;
  q.hb = 0;
  q.lr = lr;
  q.crs = crs;
  q.r2s = r2s;
  q.Tb = kTblRow * uint32_t(T);
  const int Wb = (W + 7) >> 3;
  o.frow = reinterpret_cast<DE_GLOBAL uint8_t*>(gptr(F.fwd)) + size_t(y * G + g) * Wb * 1024;
  const uint32_t fo = 16u * lane;
  const int blo = xmin >> 3, bhi = (xmax + 7) >> 3;     // blocks [blo, bhi), bhi <= Wb
  const int nblk = bhi - blo;

  // ---- H-fwd over the blocks, three stages as in k_row ----
  uint32_t t2 = 0u;
  int x = 8 * blo;
  if (nblk >= 2) {
    RowRaw R0, R1;
    RowBlk S0, S1;
    mrow_load_raw(q, R0, x, d);
    mrow_load_raw(q, R1, x + 8, d);
    row_score_raw(q, R0, S0);
    int kb = 0;                   // S0 holds block kb scored, R1 block kb+1 loaded
    for (; kb + 4 <= nblk; kb += 2, x += 16) {
      mrow_load_raw(q, R0, x + 16, d);
      row_score_raw(q, R1, S1);
      t2 = mrow_chain_fwd(S0, t2, x, xs, o.frow + size_t(x >> 3) * 1024 + fo);
      mrow_load_raw(q, R1, x + 24, d);
      row_score_raw(q, R0, S0);
      t2 = mrow_chain_fwd(S1, t2, x + 8, xs, o.frow + size_t((x + 8) >> 3) * 1024 + fo);
    }
    if (nblk - kb == 3) {
      mrow_load_raw(q, R0, x + 16, d);
      row_score_raw(q, R1, S1);
      t2 = mrow_chain_fwd(S0, t2, x, xs, o.frow + size_t(x >> 3) * 1024 + fo);
      row_score_raw(q, R0, S0);
      t2 = mrow_chain_fwd(S1, t2, x + 8, xs, o.frow + size_t((x + 8) >> 3) * 1024 + fo);
      t2 = mrow_chain_fwd(S0, t2, x + 16, xs, o.frow + size_t((x + 16) >> 3) * 1024 + fo);
    } else {
      row_score_raw(q, R1, S1);
      t2 = mrow_chain_fwd(S0, t2, x, xs, o.frow + size_t(x >> 3) * 1024 + fo);
      t2 = mrow_chain_fwd(S1, t2, x + 8, xs, o.frow + size_t((x + 8) >> 3) * 1024 + fo);
    }
  } else {
    RowRaw R0;
    RowBlk S0;
    mrow_load_raw(q, R0, x, d);
    row_score_raw(q, R0, S0);
    t2 = mrow_chain_fwd(S0, t2, x, xs, o.frow + size_t(x >> 3) * 1024 + fo);
  }

  // ---- H-bwd over the blocks, descending; a lane restarts at xe - 1 ----
  const int xl = xe - 1;
  int bb = bhi - 1;
  if (bb - P + 1 >= blo) {
    uint4 A[P], B[P];
    row_load_blocks<P>(A, o, bb);
    for (; bb - 3 * P + 1 >= blo; bb -= 2 * P) {
      row_load_blocks<P>(B, o, bb - P);
      t2 = mrow_back_blocks<P, XB>(A, t2, lr, bb, xl, o);
      row_load_blocks<P>(A, o, bb - 2 * P);
      t2 = mrow_back_blocks<P, XB>(B, t2, lr, bb - P, xl, o);
    }
    if (bb - 2 * P + 1 >= blo) {
      row_load_blocks<P>(B, o, bb - P);
      t2 = mrow_back_blocks<P, XB>(A, t2, lr, bb, xl, o);
      t2 = mrow_back_blocks<P, XB>(B, t2, lr, bb - P, xl, o);
      bb -= 2 * P;
    } else {
      t2 = mrow_back_blocks<P, XB>(A, t2, lr, bb, xl, o);
      bb -= P;
    }
  }
  for (; bb >= blo; --bb) {
    uint4 v[1];
    row_load_blocks<1>(v, o, bb);
    t2 = mrow_back_blocks<1, XB>(v, t2, lr, bb, xl, o);
  }
}

// V-fwd + V-bwd of one column and 64 planes, each plane on its own rectangle's
// y range [ys, ye); the final values are stored back into the volume. As in
// the masked row pass, the chains run over the wave's whole row span and a
// lane restarts at its own first row (forward) and last row (backward); values
// outside a plane's rectangle are never read. The scans are the full column
// pass's: buffer instructions with the row step as a scalar offset, one wide
// scalar coefficient load per chunk (the coefficient array is padded for the
// chunks that run past a column's ends), two alternating prefetch sets.
template <int G>
__global__ __launch_bounds__(64) void k_mcol(const FrameDesc* __restrict__ fp,
                                             const int4* __restrict__ rect) {
  constexpr int Dp = G * kWave;
  constexpr int U = 8;   // rows in flight
  const FrameDesc F = *fp;
  const int W = F.W, H = F.H;
  const int x = blockIdx.x / G, g = blockIdx.x - x * G;
  const int lane = int(threadIdx.x);
  const uint32_t vo = 2u * uint32_t(lane);
  const int k = g * kWave + lane;
  const int4 r = k < F.D ? rect[k] : make_int4(0, 0, 0, 0);
  const bool act = k < F.D && x >= r.z && x < r.w && r.y > r.x;
  const int ys = act ? r.x : H, ye = act ? r.y : 0;
  const int ymin = wave_min(ys), ymax = wave_max(ye);
  if (ymin >= ymax) return;   // uniform over the wave
  const int ysb = W * Dp * 2;                            // bytes per volume row
  DE_GLOBAL uint8_t* const colb =
      reinterpret_cast<DE_GLOBAL uint8_t*>(gptr(F.vol) + (size_t(x) * Dp + g * kWave));
  auto rs = [&](int y0) { return col_rsrc(colb + ptrdiff_t(y0) * ysb); };
  const DE_CONST uint32_t* ay = cptr(F.ayt) + size_t(x) * H;
  uint32_t t = 0u;

  // ---- V-fwd over rows [ymin, ymax), restart at ys ----
  {
    auto ldf = [&](uint32_t (&v)[U], int y0) {
      const __amdgpu_buffer_rsrc_t rr = rs(y0);
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = __builtin_amdgcn_raw_buffer_load_b16(rr, vo, u * ysb, 0);
    };
    auto fwd = [&](const uint32_t (&v)[U], int y0, int n) {
      const __amdgpu_buffer_rsrc_t rr = rs(y0);
      const DE_CONST uint32_t* const ap = ay + y0;
      uint32_t ab[U];
#pragma unroll
      for (int u = 0; u < U; ++u) ab[u] = ap[u];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (u < n) {
          const uint32_t tn = rf_chain_v(ab[u], t, rf_acc(ab[u], v[u]));
          t = y0 + u == ys ? v[u] << 16 : tn;
          __builtin_amdgcn_raw_buffer_store_b16(uint16_t(t >> 16), rr, vo, u * ysb, 0);
        }
      }
    };
    uint32_t A[U], B[U];
    int y = ymin;
    ldf(A, y);
    for (; y + 2 * U <= ymax; y += 2 * U) {
      ldf(B, y + U);
      fwd(A, y, U);
      ldf(A, y + 2 * U);
      fwd(B, y + U, U);
    }
    if (y + U <= ymax) {
      ldf(B, y + U);
      fwd(A, y, U);
      fwd(B, y + U, ymax - y - U);
    } else {
      fwd(A, y, ymax - y);
    }
  }

  // ---- V-bwd over rows ymax-1 .. ymin, coefficient of y+1, restart at ye-1 ----
  {
    auto ldb = [&](uint32_t (&v)[U], int ytop) {      // rows ytop, ytop-1, ..., ytop-U+1
      const __amdgpu_buffer_rsrc_t rr = rs(ytop - U + 1);
#pragma unroll
      for (int u = 0; u < U; ++u)
        v[u] = __builtin_amdgcn_raw_buffer_load_b16(rr, vo, (U - 1 - u) * ysb, 0);
    };
    const int yl = ye - 1;
    auto bwd = [&](const uint32_t (&v)[U], int ytop, int n) {
      const __amdgpu_buffer_rsrc_t rr = rs(ytop - U + 1);
      const DE_CONST uint32_t* const ap = ay + (ytop - U + 2);
      uint32_t ab[U];
#pragma unroll
      for (int u = 0; u < U; ++u) ab[u] = ap[u];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (u < n) {
          const uint32_t a2 = ab[U - 1 - u];
          const uint32_t tn = rf_chain_v(a2, t, rf_acc(a2, v[u]));
          t = ytop - u == yl ? v[u] << 16 : tn;
          __builtin_amdgcn_raw_buffer_store_b16(uint16_t(t >> 16), rr, vo, (U - 1 - u) * ysb,
                                                0);
        }
      }
    };
    uint32_t A[U], B[U];
    int y = ymax - 1;                                    // next row to process
    const int nrow = ymax - ymin;
    int done = 0;
    ldb(A, y);
    for (; done + 2 * U <= nrow; done += 2 * U, y -= 2 * U) {
      ldb(B, y - U);
      bwd(A, y, U);
      ldb(A, y - 2 * U);
      bwd(B, y - U, U);
    }
    if (done + U <= nrow) {
      ldb(B, y - U);
      bwd(A, y, U);
      bwd(B, y - U, nrow - done - U);
    } else {
      bwd(A, y, nrow - done);
    }
  }
}

// Plane k's candidates in row y: columns [xa, xb), the rectangle's x range
// cut to the row's interval; empty (xb <= xa) outside the rectangle's and the
// plane's row ranges.
__device__ __forceinline__ int2 mask_span(const int4* rect, const int4* box, const int* mx0,
                                          const int* mx1, int D, int y, int k) {
  const int4 r = rect[k], b = box[k];
  if (r.y <= r.x || y < max(r.x, b.x) || y >= min(r.y, b.y + 1)) return make_int2(0, 0);
  return make_int2(max(r.z, mx0[size_t(y) * D + k]), min(r.w, mx1[size_t(y) * D + k] + 1));
}

// The spans of every plane in row blockIdx.y, one thread a plane, into LDS
// (D <= kMaxD = the workgroup's 256 threads; planes from D on are empty).
__device__ __forceinline__ void row_spans(int2* span, const FrameDesc& F, const int4* rect,
                                          const int4* box, const int* mx0, const int* mx1) {
  const int t = threadIdx.x;
  span[t] = t < F.D ? mask_span(rect, box, mx0, mx1, F.D, blockIdx.y, t) : make_int2(0, 0);
  __syncthreads();
}

// Top-2 per pixel over the inserted planes (ascending k, strictly greater
// displaces: the packed keys, so a key of 0 changes nothing), no sub-pixel
// neighbours. The grid's rows are the image rows. A pixel's planes go 8 at a
// time: one 16-byte volume load for a group holding an inserted plane, none
// for the others (lanes are pixels 2*Dpad bytes apart, so every load touches
// a line per lane; loading a group once instead of each plane alone cuts
// the lines touched).
__global__ __launch_bounds__(256) void k_minsert(const FrameDesc* __restrict__ fp,
                                                 const int4* __restrict__ rect,
                                                 const int4* __restrict__ box,
                                                 const int* __restrict__ mx0,
                                                 const int* __restrict__ mx1) {
  static_assert(kMaxD == 256, "one thread per plane");
  __shared__ int2 span[kMaxD];
  const FrameDesc F = *fp;
  row_spans(span, F, rect, box, mx0, mx1);
  const int y = blockIdx.y;
  const int x = blockIdx.x * 256 + threadIdx.x;
  if (x >= F.W) return;
  const size_t i = size_t(y) * F.W + x;
  const uint4* pv = reinterpret_cast<const uint4*>(F.vol + i * F.Dpad);
  uint32_t p0 = 0u, p1 = 0u;
  for (int k0 = 0; k0 < F.D; k0 += 8) {                  // k0 + 7 < Dpad
    uint32_t m = 0u;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int2 sp = span[k0 + u];
      m |= uint32_t(x >= sp.x && x < sp.y) << u;
    }
    if (m) {
      const uint4 v = pv[k0 >> 3];
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t vv = (u & 1) ? (w[u >> 1] >> 16) : (w[u >> 1] & 0xFFFFu);
        const uint32_t key = (m >> u & 1u) ? pack_key(vv, k0 + u) : 0u;
        p1 = med3(p0, p1, key);
        p0 = max(p0, key);
      }
    }
  }
  reinterpret_cast<uint4*>(F.cand)[i] = make_uint4(p0, p1, 0xFFFFFFFFu, 0u);
}

// Exact reverse match over the inserted scores only (grid rows = image rows,
// as k_minsert). Right pixel xr meets plane k at left pixel xr + dmin + k.
__global__ __launch_bounds__(256) void k_mlrc(const FrameDesc* __restrict__ fp,
                                              const int4* __restrict__ rect,
                                              const int4* __restrict__ box,
                                              const int* __restrict__ mx0,
                                              const int* __restrict__ mx1) {
  __shared__ int2 span[kMaxD];
  const FrameDesc F = *fp;
  row_spans(span, F, rect, box, mx0, mx1);
  const int y = blockIdx.y;
  const int xr = blockIdx.x * 256 + threadIdx.x;
  if (xr >= F.W) return;
  uint32_t key = 0u;
  for (int k = 0; k < F.D; ++k) {
    const int x = xr + F.dmin + k;
    if (x >= F.W) break;
    const int2 sp = span[k];
    if (x >= sp.x && x < sp.y)
      key = max(key, pack_key(F.vol[(size_t(y) * F.W + x) * F.Dpad + k], k));
  }
  F.rdisp[size_t(y) * F.W + xr] = key ? int16_t(key_k(key)) : int16_t(-1);
}

hipError_t launch_masked(const FrameDesc& hf, const FrameDesc* df, const float* prior, int band,
                         const MaskBufs& m, size_t lds_row, bool lrc, hipStream_t s) {
  const int WH = hf.W * hf.H;
  hipLaunchKernelGGL(k_mask_dc, dim3((WH + 255) / 256), dim3(256), 0, s, prior, WH, m.dc);
  hipLaunchKernelGGL(k_mask_iv, dim3(hf.H), dim3(256), 0, s, df, m.dc, band, m.mx0, m.mx1);
  hipLaunchKernelGGL(k_mask_rect, dim3(hf.D), dim3(256), 0, s, df, m.mx0, m.mx1, m.box, m.rect);
  switch (hf.G) {
    case 1:
      hipLaunchKernelGGL(k_mrow<1>, dim3(hf.H), dim3(64), lds_row, s, df, m.rect);
      hipLaunchKernelGGL(k_mcol<1>, dim3(hf.W), dim3(64), 0, s, df, m.rect);
      break;
    case 2:
      hipLaunchKernelGGL(k_mrow<2>, dim3(hf.H * 2), dim3(64), lds_row, s, df, m.rect);
      hipLaunchKernelGGL(k_mcol<2>, dim3(hf.W * 2), dim3(64), 0, s, df, m.rect);
      break;
    case 3:
      hipLaunchKernelGGL(k_mrow<3>, dim3(hf.H * 3), dim3(64), lds_row, s, df, m.rect);
      hipLaunchKernelGGL(k_mcol<3>, dim3(hf.W * 3), dim3(64), 0, s, df, m.rect);
      break;
    case 4:
      hipLaunchKernelGGL(k_mrow<4>, dim3(hf.H * 4), dim3(64), lds_row, s, df, m.rect);
      hipLaunchKernelGGL(k_mcol<4>, dim3(hf.W * 4), dim3(64), 0, s, df, m.rect);
      break;
    default:
      return hipErrorInvalidValue;
  }
  const dim3 rows((hf.W + 255) / 256, hf.H);
  hipLaunchKernelGGL(k_minsert, rows, dim3(256), 0, s, df, m.rect, m.box, m.mx0, m.mx1);
  if (lrc) hipLaunchKernelGGL(k_mlrc, rows, dim3(256), 0, s, df, m.rect, m.box, m.mx0, m.mx1);
  return hipGetLastError();
}

}  // namespace de
