// Device helpers shared by the dense stage kernels (de_dense.hip, de_row.hip,
// de_col.hip): the offset-form filter steps, the row passes' staged score
// table and record operands, the volume-row transposed stores, compile-time
// unrolling, the accumulation-register moves and the buffer resource of a
// column. See de_common.h for the offset representation.
#ifndef DOUBLEEYE_DE_KERN_H
#define DOUBLEEYE_DE_KERN_H

#include <hip/hip_runtime.h>
#include <stdint.h>

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

// ---------------------------------------------------------------------------
// Row-pass pieces shared by the streaming row pass (k_rows) and the
// prior-masked one (k_mrow). Both stage the left record in its packed form
// z = kTblRow*L | 2a << 16, w = 65536 - 2a (the intensity difference masks z's
// low half first: v_sad_u16 would add the high halves' difference too).

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

// Stage the score table for the row passes: entry (a, h) = score of census
// distance h, intensity difference a <= T, at i = a * (kTblRow / 2) + h; the
// frame's table holds it at (h << sc_shift) + a. Threads tid, tid + nt, ...
// issue their loads kStage at a time before storing any (a load's latency is
// paid once per group, not once per entry).
constexpr int kStage = 8;
__device__ __forceinline__ void stage_score_table(uint16_t* tbl, const DE_CONST uint16_t* sc,
                                                  int ntt, int nb, int sc_shift, int tid,
                                                  int nt) {
  for (int i0 = tid; i0 < ntt; i0 += kStage * nt) {
    uint16_t v[kStage];
    bool ok[kStage];
#pragma unroll
    for (int u = 0; u < kStage; ++u) {
      const int i = i0 + u * nt;
      const int a = i / int(kTblRow / 2), h = i - a * int(kTblRow / 2);
      ok[u] = i < ntt && h <= nb;
      v[u] = sc[ok[u] ? (h << sc_shift) + a : 0];   // entry 0 stands in: loads unconditional
    }
#pragma unroll
    for (int u = 0; u < kStage; ++u)
      if (i0 + u * nt < ntt) tbl[i0 + u * nt] = ok[u] ? v[u] : uint16_t(kOff);
  }
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

__device__ __forceinline__ __amdgpu_buffer_rsrc_t col_rsrc(DE_GLOBAL uint8_t* p) {
  // raw buffer (stride 0); word 3 is the gfx9 data-format field of a raw
  // dword buffer; no range check is relied on (every access is in the volume)
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(p), 0, 0x7FFFFFFF,
                                           0x00020000);
}

// Device LDS per workgroup and CU count (queried once).
inline void device_shape(int* lds, int* cus) {
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

}  // namespace de

#endif
