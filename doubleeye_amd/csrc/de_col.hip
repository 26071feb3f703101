// The column pass of the dense path (stage 3: V-fwd + V-bwd + exact top-2 with
// the sub-pixel neighbours) on gfx950, with its launcher.
//
// Reference semantics (paths relative to the reference repository root):
//   int16 filter    core/tools/de_dense.cpp:411-430, 485-508 (V-fwd, V-bwd)
//   top-2 insert    core/tools/de_dense.cpp:1249-1294 (ascending k, strictly
//                   greater displaces) with the sub-pixel neighbour capture
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <utility>

#include "de_batch.h"
#include "de_common.h"
#include "de_kern.h"

namespace de {

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
// ACC: the register rows live in AGPRs (v_accvgpr_write / read, two rows a
// word; the 128 AGPRs a wave gets at two waves a SIMD hold 256 rows, the rest
// sit in VGPRs), which lets NRB = 5 or 6 blocks (320 or 384 rows) stay on
// chip at two waves a SIMD.
// NLR: V-fwd rows kept in the wave's LDS region (dynamic LDS, NLR * 128 bytes
// a wave), the NLR rows just below the register rows; 0 or kWave (one V-bwd
// block). Every frame of the launch then has H >= 64 * NRB + NLR + 2.
template <int G, int U, bool STORE, int NRB, bool ACC = false, int NLR = 0>
__global__ __launch_bounds__(64 * G) __attribute__((amdgpu_waves_per_eu(ACC ? 2 : 1)))
void k_col(BatchView b, int base) {
  constexpr int NWL = (G + 1) / 2;           // LDS words per lane per row
  constexpr int RS = kWave * NWL + 1;        // padded row stride (words)
  constexpr int RPW = (kWave + G - 1) / G;   // rows per wave in the transposed phase
  constexpr int S = G == 1 ? 1 : (G == 4 ? 4 : 2);   // lanes per row
  constexpr int NW = kWave * NWL;            // words per row
  constexpr int WPS = NW / S;                // words per lane
  constexpr int NRR = kWave * NRB;           // rows kept in registers
  static_assert(RPW * S <= kWave && NW % S == 0, "transposed phase shape");
  static_assert(kWave % (2 * U) == 0 && 2 * U <= kVolPadRows, "prefetch depth");
  static_assert(NLR == 0 || (NLR == kWave && NRB > 0 && !STORE), "LDS rows: one block");
  __shared__ uint32_t rb[kWave * RS];
  extern __shared__ uint16_t lrow_all[];                 // NLR rows x 64 lanes a wave
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
  const int ylds = yreg - NLR;                           // first LDS row
  uint16_t* const lrow = lrow_all + (threadIdx.x >> 6) * (NLR * kWave) + lane;
  uint32_t rr[NRB > 0 ? NRR / 2 : 1];

  // ---- V-fwd over rows 1 .. yend-1 in the volume, U rows in flight ----
  auto ldf = [&](uint32_t (&v)[U], int y0) {             // rows y0 .. y0+U-1
    const __amdgpu_buffer_rsrc_t rr0 = rs(y0);
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = __builtin_amdgcn_raw_buffer_load_b16(rr0, vo, u * ysb, 0);
  };
  // coefficients of rows y0 .. y0+U-1: one wide scalar load, issued a chunk
  // ahead of its use like the rows themselves (a scalar load waited on at
  // once costs the whole memory latency every U rows)
  auto ldc = [&](uint32_t (&c)[U], int y0) { __builtin_memcpy(c, ay + y0, sizeof(c)); };
  {
    auto fwd = [&](const uint32_t (&v)[U], const uint32_t (&ab)[U], int y0, int n) {
      const __amdgpu_buffer_rsrc_t rr0 = rs(y0);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (u < n) {
          const uint32_t a2 = ab[u];
          t = rf_chain(a2, t, rf_acc(a2, v[u]));
          __builtin_amdgcn_raw_buffer_store_b16(uint16_t(t >> 16), rr0, vo, u * ysb, 0);
        }
      }
    };
    const int yend = NRB > 0 ? ylds : H;
    uint32_t A[U], B[U], cA[U], cB[U];
    int y = 1;
    ldf(A, y);
    ldc(cA, y);
    for (; y + 2 * U <= yend; y += 2 * U) {
      ldf(B, y + U);
      ldc(cB, y + U);
      fwd(A, cA, y, U);
      ldf(A, y + 2 * U);
      ldc(cA, y + 2 * U);
      fwd(B, cB, y + U, U);
    }
    if (y + U <= yend) {
      ldf(B, y + U);
      ldc(cB, y + U);
      fwd(A, cA, y, U);
      fwd(B, cB, y + U, yend - y - U);
    } else {
      fwd(A, cA, y, yend - y);
    }
  }
  if constexpr (NRB > 0) {
    // ---- V-fwd over the LDS rows ylds .. yreg-1 and the register rows
    // yreg .. H-2, then row H-1 ----
    constexpr int NC = (NLR + NRR) / U;                  // chunks
    uint32_t V[2][U], AB[2][U];
    ldf(V[0], ylds);
    ldc(AB[0], ylds);
    uint32_t hold = 0u;                                  // odd row's half, waiting for its pair
    static_for<NC>([&](auto CC) __attribute__((always_inline)) {
      constexpr int c = decltype(CC)::value;
      const int y0 = ylds + c * U;
      ldf(V[(c + 1) & 1], c + 1 < NC ? y0 + U : H - 1);  // the last brings row H-1
      if constexpr (c + 1 < NC) ldc(AB[(c + 1) & 1], y0 + U);
      __builtin_amdgcn_sched_barrier(0);                 // the loads stay ahead of this chunk
      static_for<U>([&](auto UU) __attribute__((always_inline)) {
        constexpr int u = decltype(UU)::value;
        const uint32_t a2 = AB[c & 1][u];
        t = rf_chain(a2, t, rf_acc(a2, V[c & 1][u]));
        constexpr int i = c * U + u - NLR;               // register row index
        if constexpr (i < 0) {
          lrow[(c * U + u) * kWave] = uint16_t(t >> 16);   // LDS row ylds + c*U + u
        } else {
          constexpr int j = NRR - 1 - i;                 // row y0 + u = H-2-j; odd j first
          if constexpr (j & 1) {
            hold = t & 0xFFFF0000u;
          } else {
            const uint32_t w = hold | (t >> 16);
            if constexpr (ACC) rr[j >> 1] = to_acc(w);
            else rr[j >> 1] = w;
          }
        }
      });
    });
    const uint32_t a2 = ay[H - 1];
    t = rf_chain(a2, t, rf_acc(a2, V[NC & 1][0]));
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
  // coefficients of the chunk of rows ytop .. ytop-U+1: row ytop-u uses the
  // coefficient of row ytop-u+1 = c[U-1-u] (one wide scalar load, issued a
  // chunk ahead; chunks run down by U rows, through the register blocks and on
  // into the volume's, so a chunk's successor is ytop - U; reads below row 0
  // stay in the plane's padding)
  auto ldcb = [&](uint32_t (&c)[U], int ytop) {
    __builtin_memcpy(c, ay + (ytop - U + 2), sizeof(c));
  };
  auto bwd = [&](const uint32_t (&v)[U], const uint32_t (&ab)[U], int ytop, int u0, int n) {
    const __amdgpu_buffer_rsrc_t rr0 = rs(ytop - U + 1);
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
        // keys built by one byte permute each: 1 << 24 | v' << 8 | 255 - k
        // (the same order as the packed keys; the top byte marks a key, so no
        // key is 0), then turned into the packed form. Word j holds k = klo
        // and klo + 64, klo = ((j >> 6) << 7) + (j & 63), consecutive within a
        // lane's share: the constant word cw (255 - klo in byte 0, 191 - klo
        // in byte 1) steps down by 0x101 a word.
        const int j0 = sp * WPS;
        const int klo0 = ((j0 >> 6) << 7) + (j0 & 63);
        uint32_t cw = 0x01000000u | uint32_t(191 - klo0) << 8 | uint32_t(255 - klo0);
#pragma unroll 8
        for (int jj = 0; jj < WPS; ++jj) {
          const uint32_t wv = row[jj];
          const uint32_t a = __builtin_amdgcn_perm(cw, wv, 0x07010004u);   // cw.b0 wv.b0 wv.b1 cw.b3
          const uint32_t c = __builtin_amdgcn_perm(cw, wv, 0x07030205u);   // cw.b1 wv.b2 wv.b3 cw.b3
          cw -= 0x101u;
          p1 = med3(p0, p1, a);
          p0 = max(p0, a);
          p1 = med3(p0, p1, c);
          p0 = max(p0, c);
        }
        p0 -= 0xFFFF00u;                                 // to (v' + 1) << 8 | 255 - k
        p1 -= 0xFFFF00u;
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
  uint32_t cA[U], cB[U];                                // coefficient chunks in flight
  uint32_t A[U], B[U];
  const int yb0 = H - 2 - NRR - NLR;                     // first block from the volume
  if constexpr (NRB > 0) {
    // the register blocks: block k holds rows H-2-64k .. H-65-64k, slots 64k ..;
    // then the LDS block (rows yreg-1 .. ylds) when NLR > 0
    constexpr int NBK = NRB + (NLR > 0 ? 1 : 0);
    uint32_t CR[2][U];
    ldcb(CR[0], H - 2);
    static_for<NBK>([&](auto KK) __attribute__((always_inline)) {
      constexpr int k = decltype(KK)::value;
      const int yb = H - 2 - kWave * k;
      if constexpr (k == NRB) ldb(A, yb0);               // in flight over the LDS block
      static_for<kWave / U>([&](auto QQ) __attribute__((always_inline)) {
        constexpr int q = decltype(QQ)::value;
        constexpr int n = k * (kWave / U) + q;           // chunk number from the top
        const int ytop = yb - q * U;
        const __amdgpu_buffer_rsrc_t rr0 = rs(ytop - U + 1);
        ldcb(CR[(n + 1) & 1], ytop - U);                 // the next chunk's (the last
                                                         // one's: the volume's first)
        __builtin_amdgcn_sched_barrier(0);               // issued before this chunk's work
        const uint32_t (&ab)[U] = CR[n & 1];
        if constexpr (k < NRB) {
          static_for<U / 2>([&](auto PP) __attribute__((always_inline)) {
            constexpr int u = 2 * decltype(PP)::value;
            constexpr int j = kWave * k + q * U + u;     // even: rows j, j + 1 share a word
            uint32_t w;
            if constexpr (ACC) w = from_acc(rr[j >> 1]);
            else w = rr[j >> 1];
            bwd1(w & 0xFFFFu, ab[U - 1 - u], rr0, u, q * U + u);
            bwd1(w >> 16, ab[U - 2 - u], rr0, u + 1, q * U + u + 1);
          });
        } else {
          static_for<U>([&](auto UU) __attribute__((always_inline)) {
            constexpr int u = decltype(UU)::value;
            constexpr int li = NLR - 1 - (q * U + u);    // row ytop - u = ylds + li
            bwd1(uint32_t(lrow[li * kWave]), ab[U - 1 - u], rr0, u, q * U + u);
          });
        }
      });
      top2_block(yb, kWave);
    });
    constexpr int nlast = NBK * (kWave / U);             // chunks done: the next is in
#pragma unroll                                           // CR[nlast & 1]
    for (int i = 0; i < U; ++i) cA[i] = CR[nlast & 1][i];
  } else {
    ldcb(cA, H - 2);
  }
  if constexpr (NLR == 0) ldb(A, yb0);
  for (int yb = yb0; yb >= 0; yb -= kWave) {            // block: rows yb, yb-1, ..., yb-63
    const int nrow = min(kWave, yb + 1);
    int u0 = 0;
    for (; u0 + 2 * U <= nrow; u0 += 2 * U) {
      ldb(B, yb - u0 - U);
      ldcb(cB, yb - u0 - U);
      bwd(A, cA, yb - u0, u0, U);
      ldb(A, yb - u0 - 2 * U);
      ldcb(cA, yb - u0 - 2 * U);
      bwd(B, cB, yb - u0 - U, u0 + U, U);
    }
    if (u0 < nrow) {                                     // last block: fewer than 2U rows left
      if (u0 + U <= nrow) {
        ldb(B, yb - u0 - U);
        ldcb(cB, yb - u0 - U);
        bwd(A, cA, yb - u0, u0, U);
        bwd(B, cB, yb - u0 - U, u0 + U, nrow - u0 - U);
      } else {
        bwd(A, cA, yb - u0, u0, nrow - u0);
      }
    }
    top2_block(yb, nrow);
  }
}

// ---------------------------------------------------------------------------
// Host launcher.

// Column pass: rows in flight.
const int kColVariants[kNumColVariants] = {8, 16, 32};

template <int G, bool STORE, int NRB>
static void launch_col_gn(const BatchView& b, int base, int items, int variant, hipStream_t s) {
  switch (variant) {
    case 0: hipLaunchKernelGGL((k_col<G, 8, STORE, NRB>), dim3(items), dim3(64 * G), 0, s, b, base); break;
    case 1: hipLaunchKernelGGL((k_col<G, 16, STORE, NRB>), dim3(items), dim3(64 * G), 0, s, b, base); break;
    default: hipLaunchKernelGGL((k_col<G, 32, STORE, NRB>), dim3(items), dim3(64 * G), 0, s, b, base); break;
  }
}

// Register blocks by the frames' height and the launch's size. A launch of at
// most two waves per SIMD keeps 6 register blocks (384 rows; 119 VGPRs + 128
// AGPRs) and one LDS block (64 rows, 8 KB a wave beside the 16.6 KB top-2
// tile of the workgroup: four workgroups of two waves a CU still fit) when
// every frame has H >= 450 and two or more disparity groups; 5 register
// blocks and the LDS block from H >= 386; 5 register blocks alone from
// H >= 322 (at the default 16 rows in flight), else 3 blocks (192 rows, ~130
// more VGPRs) when every frame has H >= 194; a larger launch 2 blocks (128
// rows), which leave room for more waves per SIMD; frames under 130 rows none.
// (One disparity group keeps no LDS block: a one-wave workgroup's 16.6 KB tile
// plus 8 KB would cut the CU to six waves.) Measured at 848x480x128 (1696
// waves): column pass 66 us without register blocks, 62 with 2, 60 with 3,
// 53.7 with 5, 51.4 with 5 + LDS, 47.0 with 6 + LDS (31 rows left to the
// volume); four frames in flight 0.141 / 0.137 ms a frame with 2 / 3 blocks.
// The 64-pair batch ran 10.9 ms with 2 blocks, 11.9 with 3. Four blocks in
// VGPRs (256 VGPRs, spills) took 76 us.
template <int G, bool STORE>
static void launch_col_g(const BatchView& b, int base, int items, int variant, int min_h,
                         hipStream_t s) {
  int lds = 0, cus = 1;
  device_shape(&lds, &cus);
  const bool small = size_t(items) * G <= size_t(8) * cus;
  if constexpr (STORE) {
    // the volume is written back anyway (tests, the exact reverse match)
    launch_col_gn<G, STORE, 0>(b, base, items, variant, s);
  } else {
    if (G >= 2 && small && min_h >= 7 * kWave + 2 && variant == 1)
      hipLaunchKernelGGL((k_col<G, 16, STORE, 6, true, kWave>), dim3(items), dim3(64 * G),
                         size_t(G) * kWave * kWave * 2, s, b, base);
    else if (G >= 2 && small && min_h >= 6 * kWave + 2 && variant == 1)
      hipLaunchKernelGGL((k_col<G, 16, STORE, 5, true, kWave>), dim3(items), dim3(64 * G),
                         size_t(G) * kWave * kWave * 2, s, b, base);
    else if (small && min_h >= 5 * kWave + 2 && variant == 1)
      hipLaunchKernelGGL((k_col<G, 16, STORE, 5, true>), dim3(items), dim3(64 * G), 0, s, b, base);
    else if (small && min_h >= 3 * kWave + 2) launch_col_gn<G, STORE, 3>(b, base, items, variant, s);
    else if (min_h >= 2 * kWave + 2) launch_col_gn<G, STORE, 2>(b, base, items, variant, s);
    else launch_col_gn<G, STORE, 0>(b, base, items, variant, s);
  }
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

}  // namespace de
