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
// consecutive k, so every volume access is one 128-byte coalesced segment.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "de_batch.h"
#include "de_common.h"

namespace de {

// 2*f' = a2 * hi16(t2prev) + acc, with a2 = 2a wave-uniform (SGPR) and
// acc = (65536 - a2) * c' + 32768 computed off the dependency chain.
__device__ __forceinline__ uint32_t rf_chain(uint32_t a2, uint32_t t2prev, uint32_t acc) {
  uint32_t r;
  asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[0,1,0,0]" : "=v"(r) : "s"(a2), "v"(t2prev), "v"(acc));
  return r;
}
__device__ __forceinline__ uint32_t rf_acc(uint32_t a2, uint32_t c) {
  return __umul24(65536u - a2, c) + 32768u;
}
__device__ __forceinline__ uint16_t hi16(uint32_t t2) { return uint16_t(t2 >> 16); }

__device__ __forceinline__ int local_item(const BatchView& b, int stage, int item, int f) {
  return item - b.start[stage * (b.n + 1) + f];
}

// ---------------------------------------------------------------------------
// Stage 1: census of both images, left record, transposed vertical coefficients.
// Tile of 64 x 8 pixels staged in LDS with its census halo; 256 threads, two
// pixels each.
template <int HW, int HH>
__global__ __launch_bounds__(256) void k_prep(BatchView b) {
  constexpr int TW = kPrepTileW + 2 * HW;
  constexpr int TH = kPrepTileH + 2 * HH;
  __shared__ uint8_t sl[TH][TW];
  __shared__ uint8_t sr[TH][TW];
  __shared__ uint16_t srf[256];
  const int item = blockIdx.x;
  const int f = find_frame(b, kStagePrep, item);
  const FrameDesc& F = b.frames[f];
  const int local = local_item(b, kStagePrep, item, f);
  const int W = F.W, H = F.H;
  const int ntx = (W + kPrepTileW - 1) / kPrepTileW;
  const int x0 = (local % ntx) * kPrepTileW, y0 = (local / ntx) * kPrepTileH;
  srf[threadIdx.x] = F.rf[threadIdx.x];
  for (int i = threadIdx.x; i < TW * TH; i += 256) {
    const int ty = i / TW, tx = i - ty * TW;
    const int gx = x0 - HW + tx, gy = y0 - HH + ty;
    const bool in = gx >= 0 && gx < W && gy >= 0 && gy < H;
    sl[ty][tx] = in ? F.L[size_t(gy) * W + gx] : 0;
    sr[ty][tx] = in ? F.R[size_t(gy) * W + gx] : 0;
  }
  __syncthreads();
  const int lx = threadIdx.x & 63;
  for (int r = threadIdx.x >> 6; r < kPrepTileH; r += 4) {
    const int x = x0 + lx, y = y0 + r;
    if (x >= W || y >= H) continue;
    const int cy = r + HH, cx = lx + HW;
    uint64_t vl = 0, vr = 0;
    if (x >= HW && x < W - HW && y >= HH && y < H - HH) {
      const uint8_t ml = sl[cy][cx], mr = sr[cy][cx];
      int bit = 0;
#pragma unroll
      for (int dy = -HH; dy <= HH; ++dy)
#pragma unroll
        for (int dx = -HW; dx <= HW; ++dx) {
          if (dx == 0 && dy == 0) continue;
          vl |= uint64_t(sl[cy + dy][cx + dx] < ml) << bit;
          vr |= uint64_t(sr[cy + dy][cx + dx] < mr) << bit;
          ++bit;
        }
    }
    const int lc = sl[cy][cx];
    const uint32_t ax = x ? srf[abs(lc - int(sl[cy][cx - 1]))] : srf[0];
    const uint32_t ay = y ? srf[abs(lc - int(sl[cy - 1][cx]))] : srf[0];
    const size_t i = size_t(y) * W + x;
    uint4 rec;
    rec.x = uint32_t(vl);
    rec.y = uint32_t(vl >> 32);
    rec.z = uint32_t(2 * lc) | ((2u * ax) << 16);
    rec.w = 0;
    reinterpret_cast<uint4*>(F.lrec)[i] = rec;
    F.cr[i] = vr;
    F.ayt[size_t(x) * H + y] = 2u * ay;
  }
}

// ---------------------------------------------------------------------------
// Stage 2: graded score + H-fwd + H-bwd for one row and 64 disparities.
// Lane = disparity; the left record of pixel x is wave-uniform (scalar loads),
// the right census/intensity at x - d comes from the LDS-staged row.
// Dynamic LDS: score table | right census row (u64) | 2*R row (u16).
__device__ __forceinline__ uint32_t score_at(const uint8_t* smem, const uint64_t* crs,
                                             const uint16_t* r2s, uint4 lrec, int j,
                                             int sh, uint32_t T2) {
  const uint64_t cr = crs[j];
  const uint32_t h = __popc(lrec.x ^ uint32_t(cr)) + __popc(lrec.y ^ uint32_t(cr >> 32));
  const uint32_t ad2 = min(__usad(lrec.z & 0xFFFFu, uint32_t(r2s[j]), 0u), T2);
  return *reinterpret_cast<const uint16_t*>(smem + ((h << sh) + ad2));
}

__global__ __launch_bounds__(64) void k_row(BatchView b) {
  extern __shared__ __align__(16) uint8_t smem[];
  const int item = blockIdx.x;
  const int f = find_frame(b, kStageRow, item);
  const FrameDesc& F = b.frames[f];
  const int local = local_item(b, kStageRow, item, f);
  const int W = F.W, H = F.H, Dp = F.Dpad, G = F.G;
  const int y = local / G, g = local - y * G;
  const int lane = threadIdx.x;
  const int k = g * kWave + lane;
  const int d = F.dmin + k;
  uint16_t* vcol = F.vol + size_t(y) * W * Dp + k;      // element x at vcol[x*Dp]
  const int xlo = F.hw + F.dmin + g * kWave;             // lane 0 valid from here
  const int xend = W - F.hw;                             // left pixel valid below
  if (y < F.hh || y >= H - F.hh || xend <= xlo) {
    // every score of this row is zero -> the filtered row is zero (v' = kOff)
    for (int x = 0; x < W; ++x) vcol[size_t(x) * Dp] = uint16_t(kOff);
    return;
  }
  const int nb = (2 * F.hw + 1) * (2 * F.hh + 1) - 1;
  const int ntbl = (nb + 1) << F.sc_shift;
  uint64_t* crs = reinterpret_cast<uint64_t*>(smem + ((ntbl * 2 + 15) & ~15));
  uint16_t* r2s = reinterpret_cast<uint16_t*>(crs + W);
  {
    uint16_t* tbl = reinterpret_cast<uint16_t*>(smem);
    for (int i = lane; i < ntbl; i += kWave) tbl[i] = F.sc[i];
    const uint64_t* crg = F.cr + size_t(y) * W;
    const uint8_t* Rg = F.R + size_t(y) * W;
    for (int x = lane; x < W; x += kWave) { crs[x] = crg[x]; r2s[x] = uint16_t(2 * Rg[x]); }
  }
  __syncthreads();
  const uint4* __restrict__ lr = reinterpret_cast<const uint4*>(F.lrec) + size_t(y) * W;
  const int sh = F.sc_shift + 1;
  const uint32_t T2 = 2u * uint32_t(F.ad_T);
  const int xa = xlo;                                    // xlo < xend here
  const int xb = max(xa, min(xlo + kWave - 1, xend));

  // ---- H-fwd: x in [0, xa) every score is zero, so f' stays kOff ----
  for (int x = 0; x < xa; ++x) vcol[size_t(x) * Dp] = uint16_t(kOff);
  uint32_t t2 = uint32_t(kOff) << 16;
  for (int x = xa; x < xb; ++x) {                        // ragged band: per-lane window
    const uint4 rec = lr[x];
    const bool ok = x >= F.hw + d;
    const uint32_t c = ok ? score_at(smem, crs, r2s, rec, ok ? x - d : 0, sh, T2) : uint32_t(kOff);
    const uint32_t a2 = __builtin_amdgcn_readfirstlane(rec.z >> 16);
    t2 = rf_chain(a2, t2, rf_acc(a2, c));
    vcol[size_t(x) * Dp] = hi16(t2);
  }
#pragma unroll 4
  for (int x = xb; x < xend; ++x) {                      // every lane inside its window
    const uint4 rec = lr[x];
    const uint32_t c = score_at(smem, crs, r2s, rec, x - d, sh, T2);
    const uint32_t a2 = __builtin_amdgcn_readfirstlane(rec.z >> 16);
    t2 = rf_chain(a2, t2, rf_acc(a2, c));
    vcol[size_t(x) * Dp] = hi16(t2);
  }
  for (int x = xend; x < W; ++x) {                       // right border: zero scores
    const uint32_t a2 = __builtin_amdgcn_readfirstlane(lr[x].z >> 16);
    t2 = rf_chain(a2, t2, rf_acc(a2, uint32_t(kOff)));
    vcol[size_t(x) * Dp] = hi16(t2);
  }

  // ---- H-bwd: g'[W-1] = f'[W-1]; coefficient of x+1 ----
  constexpr int U = 8;
  int x = W - 2;
  uint32_t cur[U];
  // software pipeline: the U forward values of the next chunk are in flight
  // while the current chunk runs its chain
  const int xs = xa + U - 1;                             // chunks cover [xa, x]
  if (x >= xs) {
#pragma unroll
    for (int u = 0; u < U; ++u) cur[u] = vcol[size_t(x - u) * Dp];
  }
  for (; x >= xs; x -= U) {
    uint32_t nxt[U];
    const bool more = x - U >= xs;
    if (more) {
#pragma unroll
      for (int u = 0; u < U; ++u) nxt[u] = vcol[size_t(x - U - u) * Dp];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t a2 = __builtin_amdgcn_readfirstlane(lr[x - u + 1].z >> 16);
      t2 = rf_chain(a2, t2, rf_acc(a2, cur[u]));
      vcol[size_t(x - u) * Dp] = hi16(t2);
    }
    if (more) {
#pragma unroll
      for (int u = 0; u < U; ++u) cur[u] = nxt[u];
    }
  }
  for (; x >= xa; --x) {
    const uint32_t c = vcol[size_t(x) * Dp];
    const uint32_t a2 = __builtin_amdgcn_readfirstlane(lr[x + 1].z >> 16);
    t2 = rf_chain(a2, t2, rf_acc(a2, c));
    vcol[size_t(x) * Dp] = hi16(t2);
  }
  for (; x >= 0; --x) {                                  // forward value is kOff here
    const uint32_t a2 = __builtin_amdgcn_readfirstlane(lr[x + 1].z >> 16);
    t2 = rf_chain(a2, t2, rf_acc(a2, uint32_t(kOff)));
    vcol[size_t(x) * Dp] = hi16(t2);
  }
}

// ---------------------------------------------------------------------------
// Stage 3: V-fwd + V-bwd + exact top-2 with sub-pixel neighbours, one wave per
// column x holding all G disparity groups (G chains per lane). The V-bwd values
// never go back to memory; the per-pixel output is one 16-byte record.
__device__ __forceinline__ void top2_insert(uint32_t& p0, uint32_t& p1, uint32_t key) {
  // distinct keys: new best = max, new runner-up = median of three
  p1 = max(p1, min(p0, key));
  p0 = max(p0, key);
}

__device__ __forceinline__ void top2_merge(uint32_t& p0, uint32_t& p1, uint32_t o0, uint32_t o1) {
  const uint32_t n1 = max(min(p0, o0), max(p1, o1));
  p0 = max(p0, o0);
  p1 = n1;
}

template <int GM>
__global__ __launch_bounds__(64) void k_col(BatchView b) {
  const int item = blockIdx.x;
  const int f = find_frame(b, kStageCol, item);
  const FrameDesc& F = b.frames[f];
  const int x = local_item(b, kStageCol, item, f);
  const int W = F.W, H = F.H, Dp = F.Dpad, G = F.G;
  const int lane = threadIdx.x;
  uint16_t* vx = F.vol + size_t(x) * Dp + lane;          // (y, g) at vx[y*ys + g*64]
  const size_t ys = size_t(W) * Dp;
  const uint32_t* __restrict__ ay = F.ayt + size_t(x) * H;
  uint32_t t[GM];
#pragma unroll
  for (int g = 0; g < GM; ++g) t[g] = g < G ? uint32_t(vx[g * kWave]) << 16 : 0u;

  // ---- V-fwd ----
  for (int y = 1; y < H; ++y) {
    const uint32_t a2 = __builtin_amdgcn_readfirstlane(ay[y]);
    uint16_t* p = vx + size_t(y) * ys;
#pragma unroll
    for (int g = 0; g < GM; ++g) {
      if (g < G) {
        t[g] = rf_chain(a2, t[g], rf_acc(a2, p[g * kWave]));
        p[g * kWave] = hi16(t[g]);
      }
    }
  }

  // ---- V-bwd + top-2 ----
  const int kmax = (x >= F.hw + F.dmin && x < W - F.hw) ? min(F.D - 1, x - F.hw - F.dmin) : -1;
  uint4* rec = reinterpret_cast<uint4*>(F.cand) + x;    // row y at rec[y*W]
  if (lane == 0) rec[size_t(H - 1) * W] = make_uint4(0u, 0u, 0xFFFFFFFFu, 0u);
  for (int y = H - 2; y >= 0; --y) {
    const uint32_t a2 = __builtin_amdgcn_readfirstlane(ay[y + 1]);
    const uint16_t* p = vx + size_t(y) * ys;
    uint32_t v[GM];
#pragma unroll
    for (int g = 0; g < GM; ++g) {
      if (g < G) {
        t[g] = rf_chain(a2, t[g], rf_acc(a2, p[g * kWave]));
        v[g] = t[g] >> 16;
        if (F.store_vol) const_cast<uint16_t*>(p)[g * kWave] = uint16_t(v[g]);
      } else {
        v[g] = 0;
      }
    }
    if (y < F.hh || y >= H - F.hh || kmax < 0) {
      if (lane == 0) rec[size_t(y) * W] = make_uint4(0u, 0u, 0xFFFFFFFFu, 0u);
      continue;
    }
    uint32_t p0 = 0, p1 = 0;
#pragma unroll
    for (int g = 0; g < GM; ++g) {
      const int kk = g * kWave + lane;
      if (g < G && kk <= kmax) top2_insert(p0, p1, pack_key(v[g], kk));
    }
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t o0 = __shfl_xor(p0, off);
      const uint32_t o1 = __shfl_xor(p1, off);
      top2_merge(p0, p1, o0, o1);
    }
    p0 = __builtin_amdgcn_readfirstlane(p0);
    p1 = __builtin_amdgcn_readfirstlane(p1);
    uint32_t nb = 0xFFFFFFFFu;
    if (p0) {
      const int ks = key_k(p0);
      uint32_t nm = 0xFFFFu, np = 0xFFFFu;
      if (ks >= 1) {
        const int km = ks - 1, gm = km >> 6;
        uint32_t s = 0;
#pragma unroll
        for (int g = 0; g < GM; ++g) s = (g == gm) ? v[g] : s;
        nm = __builtin_amdgcn_readlane(s, km & 63);
      }
      if (ks + 1 <= kmax) {
        const int kp = ks + 1, gp = kp >> 6;
        uint32_t s = 0;
#pragma unroll
        for (int g = 0; g < GM; ++g) s = (g == gp) ? v[g] : s;
        np = __builtin_amdgcn_readlane(s, kp & 63);
      }
      nb = nm | (np << 16);
    }
    if (lane == 0) rec[size_t(y) * W] = make_uint4(p0, p1, nb, 0u);
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
      const int cy = r + HH, cx = lx + HW;
      const uint8_t m = s[cy][cx];
      int bit = 0;
#pragma unroll
      for (int dy = -HH; dy <= HH; ++dy)
#pragma unroll
        for (int dx = -HW; dx <= HW; ++dx) {
          if (dx == 0 && dy == 0) continue;
          v |= uint64_t(s[cy + dy][cx + dx] < m) << bit;
          ++bit;
        }
    }
    out[size_t(y) * W + x] = v;
  }
}

// ---------------------------------------------------------------------------
// Host launchers (called from the C-ABI translation unit).
size_t row_lds_bytes(int W, int nbits, int sc_shift) {
  const int ntbl = (nbits + 1) << sc_shift;
  return size_t((ntbl * 2 + 15) & ~15) + size_t(W) * 8 + size_t(W) * 2;
}

hipError_t launch_prep(const BatchView& b, int items, int hw, int hh, hipStream_t s) {
  if (hw == 3 && hh == 3) hipLaunchKernelGGL((k_prep<3, 3>), dim3(items), dim3(256), 0, s, b);
  else if (hw == 4 && hh == 3) hipLaunchKernelGGL((k_prep<4, 3>), dim3(items), dim3(256), 0, s, b);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_row(const BatchView& b, int items, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL(k_row, dim3(items), dim3(64), lds, s, b);
  return hipGetLastError();
}

hipError_t launch_col(const BatchView& b, int items, int gmax, hipStream_t s) {
  switch (gmax) {
    case 1: hipLaunchKernelGGL(k_col<1>, dim3(items), dim3(64), 0, s, b); break;
    case 2: hipLaunchKernelGGL(k_col<2>, dim3(items), dim3(64), 0, s, b); break;
    case 3: hipLaunchKernelGGL(k_col<3>, dim3(items), dim3(64), 0, s, b); break;
    case 4: hipLaunchKernelGGL(k_col<4>, dim3(items), dim3(64), 0, s, b); break;
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

}  // namespace de
