// Dense image-plane stages for gfx950: census + filter coefficients (stage
// 1), the standalone census plane, the exact reverse match and the
// prior-masked passes. The row pass is in de_row.hip, the column pass in
// de_col.hip, their shared helpers in de_kern.h.
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
#include "de_kern.h"

namespace de {

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
  const uint16_t rfv = cptr(F.rf)[threadIdx.x];
  {
    // every load of the tile is issued before the first LDS store: one memory
    // latency for the staging, not one per pass of the 256 threads
    constexpr int NTB = TWP * TH, IT = (NTB + 255) / 256;
    uint32_t vl[IT], vr[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const int i = threadIdx.x + 256 * k;
      const int ty = i / TWP, tx = i - ty * TWP;
      const int gx = x0 - HW + tx, gy = y0 - HH + ty;
      const bool in = i < NTB && gx >= 0 && gx < W && gy >= 0 && gy < H;
      const size_t gi = in ? size_t(gy) * W + gx : 0;    // pixel 0 stands in: loads unconditional
      vl[k] = Lg[gi];
      vr[k] = Rg[gi];
      vl[k] = in ? vl[k] : 0u;
      vr[k] = in ? vr[k] : 0u;
    }
    srf[threadIdx.x] = rfv;
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const int i = threadIdx.x + 256 * k;
      if (i < NTB) {
        const int ty = i / TWP, tx = i - ty * TWP;
        st[0][ty][tx] = uint8_t(vl[k]);
        st[1][ty][tx] = uint8_t(vr[k]);
      }
    }
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
