// FAST keypoint detector, census at keypoints, SAD sub-pixel refinement and the
// detect -> describe -> match pipeline on the MI355X.
//
// detect_keypoints_fast (core/src/preproc.cpp:436-545) runs as a fixed chain of
// launches with no host synchronisation, batched over frames (blockIdx.z /
// blockIdx.y = frame):
//   k_det_init     zero each frame's counters and cell counts (the FAST kernel
//                  does it when the suppression is tiled)
//   k_det_fast     FAST-9 (preproc.cpp:371-399) from a 70x10 LDS tile per
//                  64x4 pixels, Shi-Tomasi (preproc.cpp:401-427) at the
//                  corners; survivors of min_response join a candidate list
//   k_det_nms_tile suppression passes over 64x16 tiles staged in LDS with a
//                  halo of nms_radius, each iterated to a fixed point; repeated
//                  kNmsTilePasses (k_det_nms: global rounds for wider radii)
//   k_det_nms_fin  one workgroup per frame finishes any undecided candidates
//   k_det_refine   per kept candidate: texture floor (local_std_at, 7x7),
//                  sub-pixel parabola, cell counts
//   k_det_scan     exclusive scans over the cells
//   k_det_scatter  kept candidates into per-cell lists
//   k_det_select   per cell, the top per_cell by (response desc, scan order)
//
// The reference suppresses greedily: candidates in descending response order,
// each kept unless an already kept candidate lies within the square window of
// radius nms_radius. That is the lexicographically first maximal independent set
// of the "within the window" graph in priority order. The rounds compute it
// without a sort: an undecided candidate becomes kept once no neighbour in its
// window is kept or undecided with a higher priority, and is removed once a
// neighbour is kept. Transitions are monotone (undecided -> kept / removed), so
// reads that race a neighbour's transition only postpone a decision, and by
// induction on priority every decision equals the sequential one. Equal
// responses are ordered by scan position (y, then x); the reference's std::sort
// leaves their order unspecified, so results equal its output whenever equal
// responses do not compete within a window or for a cell's last places.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "de_api_internal.h"

namespace de {

namespace {

constexpr int kTileW = 64, kTileH = 4, kHalo = 3;      // FAST tile (one pixel per thread)
constexpr int kLdsW = kTileW + 2 * kHalo, kLdsH = kTileH + 2 * kHalo;
constexpr int kDetThreads = 256;
constexpr int kFinThreads = 1024;
constexpr int kFinList = 8192;      // undecided candidates the finisher holds in LDS
constexpr int kNtW = 64, kNtH = 16, kNtMaxRad = 8;      // suppression tile, widest LDS halo
constexpr int kNtLW = kNtW + 2 * kNtMaxRad, kNtLH = kNtH + 2 * kNtMaxRad;
constexpr int kNmsTilePasses = 3;   // LDS-tile passes before the finisher
constexpr int kNmsRounds = 8;       // global rounds before the finisher (radius > kNtMaxRad)
constexpr int kFlags = kNmsRounds + 1;   // [r]: some candidate undecided before pass r
constexpr int kMaxNmsRadius = 255;

// FAST ring: Bresenham circle of radius 3 from the top, clockwise.
__constant__ int8_t c_ring_x[16] = {0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3, -3, -3, -2, -1};
__constant__ int8_t c_ring_y[16] = {-3, -3, -2, -1, 0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3};

struct DetDesc {
  const uint8_t* img;
  int W, H;
  int border, thr, r, rad;   // border = max(cfg.border, 5), r = max(1, grad_window/2)
  int cell, cols, rows, keep, subpixel, ncells;
  int tiled;                 // 1: LDS-tile suppression (nms radius <= kNtMaxRad)
  float min_resp, min_std;
  // scratch
  uint8_t* state;            // W*H: 0 none / removed, 1 undecided, 2 kept
  float* resp;               // W*H: Shi-Tomasi response of candidates
  uint32_t* cand;            // candidates left to the finisher (tiled), or all candidates
  uint32_t* kept;            // pixel indices of the candidates suppression keeps
  int* cnt;                  // [0] candidates, [1] kept by suppression, [2] keypoints,
                             // [3] candidates left undecided by the tile passes
  float4* rec;               // per kept candidate: x, y, response, local_std
  int* slot;                 // per kept candidate: place in its cell's list, -1 = dropped
  int* cell_cnt;             // [ncells]
  int* cell_off;             // [ncells + 1]
  int* out_off;              // [ncells + 1]
  int* items;                // kept candidates grouped by cell
  int* tpend;                // [2][tiles]: tile still holds undecided candidates
  // outputs
  float4* out;
  int* n_out;
};

// A run of at least 9 set bits in the 16-bit circular mask m: after the three
// AND-shift steps bit i holds bits i..i+7 of the doubled mask, and one more
// AND with the mask shifted by 8 extends that to i..i+8.
__device__ __forceinline__ bool run9(uint32_t m) {
  const uint32_t d = m | (m << 16);
  uint32_t t = d & (d >> 1);
  t &= t >> 2;
  t &= t >> 4;
  t &= d >> 8;
  return (t & 0xFFFFu) != 0u;
}

// is_fast_corner over pixel access px(dx, dy) relative to the centre; the
// caller has checked 3 <= x < W-3 and 3 <= y < H-3.
template <class Px>
__device__ __forceinline__ bool fast_at(Px px, int thr) {
  const int c = px(0, 0);
  const int hi = c + thr, lo = c - thr;
  int bright = 0, dark = 0;
#pragma unroll
  for (int k = 0; k < 16; k += 4) {
    const int v = px(c_ring_x[k], c_ring_y[k]);
    if (v > hi) ++bright;
    else if (v < lo) ++dark;
  }
  if (bright < 2 && dark < 2) return false;
  uint32_t bm = 0, dm = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int v = px(c_ring_x[k], c_ring_y[k]);
    if (v > hi) bm |= 1u << k;
    else if (v < lo) dm |= 1u << k;
  }
  return run9(bm) || run9(dm);
}

// shi_tomasi_at: structure tensor of the 3x3 Sobel gradients (scaled by 1/8)
// summed over the (2r+1)^2 window, dy outer and dx inner, then its smaller
// eigenvalue. The caller has checked that the window plus one pixel is inside.
template <class Px>
__device__ __forceinline__ float shi_at(Px px, int r) {
  float sxx = 0.f, syy = 0.f, sxy = 0.f;
  for (int dy = -r; dy <= r; ++dy) {
    for (int dx = -r; dx <= r; ++dx) {
      const float gx = (float(px(dx + 1, dy - 1)) - float(px(dx - 1, dy - 1))
                        + 2.f * (float(px(dx + 1, dy)) - float(px(dx - 1, dy)))
                        + float(px(dx + 1, dy + 1)) - float(px(dx - 1, dy + 1))) * 0.125f;
      const float gy = (float(px(dx - 1, dy + 1)) - float(px(dx - 1, dy - 1))
                        + 2.f * (float(px(dx, dy + 1)) - float(px(dx, dy - 1)))
                        + float(px(dx + 1, dy + 1)) - float(px(dx + 1, dy - 1))) * 0.125f;
      sxx += gx * gx;
      syy += gy * gy;
      sxy += gx * gy;
    }
  }
  const float n = 1.f / float((2 * r + 1) * (2 * r + 1));
  const float a = sxx * n, b = syy * n, c = sxy * n;
  const float half_tr = 0.5f * (a + b);
  const float q = 0.25f * (a - b) * (a - b) + c * c;
  const float disc = sqrtf((0.f < q) ? q : 0.f);
  return half_tr - disc;
}

struct GlobalPx {
  const uint8_t* img;
  int W, x, y;
  __device__ int operator()(int dx, int dy) const {
    return img[size_t(y + dy) * W + (x + dx)];
  }
};

struct TilePx {
  const uint8_t (*t)[kLdsW];
  int tx, ty;
  __device__ int operator()(int dx, int dy) const { return t[ty + dy][tx + dx]; }
};

// shi_tomasi_at with its bounds check (0 outside), from global memory.
__device__ float shi_global(const uint8_t* img, int W, int H, int x, int y, int r) {
  if (x - r - 1 < 0 || y - r - 1 < 0 || x + r + 1 >= W || y + r + 1 >= H) return 0.f;
  return shi_at(GlobalPx{img, W, x, y}, r);
}

// local_std_at(img, x, y, 7): population std over the clipped 7x7 window, in
// double as the reference computes it.
__device__ float local_std7(const uint8_t* img, int W, int H, int x, int y) {
  const int x0 = max(0, x - 3), x1 = min(W - 1, x + 3);
  const int y0 = max(0, y - 3), y1 = min(H - 1, y + 3);
  double s1 = 0.0, s2 = 0.0;
  for (int yy = y0; yy <= y1; ++yy)
    for (int xx = x0; xx <= x1; ++xx) {
      const double v = img[size_t(yy) * W + xx];
      s1 += v;
      s2 += v * v;
    }
  const double n = double((y1 - y0 + 1) * (x1 - x0 + 1));
  const double mean = s1 / n;
  const double var = s2 / n - mean * mean;
  return var > 0.0 ? float(sqrt(var)) : 0.f;
}

// Order-preserving map of a float to uint32 (-0 and +0 map alike, as `<` sees them).
__device__ __forceinline__ uint32_t ordered_f32(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u << 1) == 0u) u = 0u;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ uint64_t lanes_below() {
  const int lane = __lane_id();
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kDetThreads) void k_det_init(const DetDesc* __restrict__ frames,
                                                          int* __restrict__ flags) {
  const DetDesc d = frames[blockIdx.y];
  const int t = blockIdx.x * kDetThreads + threadIdx.x;
  if (t < 4) d.cnt[t] = 0;
  if (t < d.ncells) d.cell_cnt[t] = 0;
  if (blockIdx.y == 0 && t < kFlags) flags[t] = (t == 0) ? 1 : 0;
}

// FAST + Shi-Tomasi over a 64x4 block of pixels, one per thread; every pixel
// of the image gets its state written. With tiled suppression the tile passes
// gather candidates from the state map, and block (0, 0) zeroes the frame's
// counters (and frame 0's the pass flags): nothing else touches them before the
// first tile pass. Otherwise candidates join a list, one atomic per block.
__global__ __launch_bounds__(kDetThreads) void k_det_fast(const DetDesc* __restrict__ frames,
                                                          int* __restrict__ flags) {
  const DetDesc d = frames[blockIdx.z];
  const int W = d.W, H = d.H;
  const int x0 = blockIdx.x * kTileW, y0 = blockIdx.y * kTileH;
  if (x0 >= W || y0 >= H) return;   // uniform over the block
  __shared__ uint8_t tile[kLdsH][kLdsW];
  __shared__ int wave_n[kDetThreads / kWave];
  __shared__ int base_sh;
  if (d.tiled && blockIdx.x == 0 && blockIdx.y == 0) {
    for (int t = threadIdx.x; t < d.ncells; t += kDetThreads) d.cell_cnt[t] = 0;
    if (threadIdx.x < 4) d.cnt[threadIdx.x] = 0;
    if (blockIdx.z == 0 && threadIdx.x < kFlags) flags[threadIdx.x] = threadIdx.x == 0 ? 1 : 0;
  }
  for (int i = threadIdx.x; i < kLdsH * kLdsW; i += kDetThreads) {
    const int ty = i / kLdsW, tx = i - ty * kLdsW;
    const int gx = min(max(x0 - kHalo + tx, 0), W - 1);
    const int gy = min(max(y0 - kHalo + ty, 0), H - 1);
    tile[ty][tx] = d.img[size_t(gy) * W + gx];
  }
  __syncthreads();
  const int lx = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const int b = d.border, r = d.r;
  const int x = x0 + lx, y = y0 + wv;
  bool is_cand = false;
  if (x < W && y < H) {
    if (x >= b && x < W - b && y >= b && y < H - b) {
      const TilePx tp{tile, lx + kHalo, wv + kHalo};
      if (fast_at(tp, d.thr)) {
        float resp;
        if (r + 1 <= kHalo) resp = shi_at(tp, r);
        else resp = shi_global(d.img, W, H, x, y, r);
        if (!(resp < d.min_resp)) {
          is_cand = true;
          d.resp[size_t(y) * W + x] = resp;
        }
      }
    }
    d.state[size_t(y) * W + x] = is_cand ? 1 : 0;
  }
  if (d.tiled) return;   // uniform over the block
  const uint64_t m = __ballot(is_cand);
  if (lx == 0) wave_n[wv] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < kDetThreads / kWave; ++w) tot += wave_n[w];
    base_sh = tot ? atomicAdd(&d.cnt[0], tot) : 0;
  }
  __syncthreads();
  if (is_cand) {
    int before = base_sh;
    for (int w = 0; w < wv; ++w) before += wave_n[w];
    d.cand[before + __popcll(m & lanes_below())] = uint32_t(y * W + x);
  }
}

// Neighbour (dy, dx) outranks a candidate of response k: a higher response, or
// an equal one earlier in scan order.
__device__ __forceinline__ bool outranks(float kn, float k, int dy, int dx) {
  return kn > k || (kn == k && (dy < 0 || (dy == 0 && dx < 0)));
}

// Decide candidate p once from global memory: 2 = kept, 0 = removed, 1 = still
// undecided. Radii up to 3 load the whole window at once.
__device__ __forceinline__ int nms_decide(const DetDesc& d, uint32_t p) {
  const int W = d.W, H = d.H, rad = d.rad;
  if (rad <= 3) {
    const int y = int(p / uint32_t(W)), x = int(p - uint32_t(y) * uint32_t(W));
    uint8_t s[7][7];
#pragma unroll
    for (int a = 0; a < 7; ++a)
#pragma unroll
      for (int b = 0; b < 7; ++b) {
        const int dy = a - 3, dx = b - 3, py = y + dy, px = x + dx;
        const bool ok = (dx | dy) != 0 && abs(dy) <= rad && abs(dx) <= rad && py >= 0 &&
                        py < H && px >= 0 && px < W;
        s[a][b] = ok ? static_cast<const volatile uint8_t*>(d.state)[size_t(py) * W + px] : 0;
      }
    bool any1 = false;
#pragma unroll
    for (int a = 0; a < 7; ++a)
#pragma unroll
      for (int b = 0; b < 7; ++b) {
        if (s[a][b] == 2) return 0;
        any1 = any1 || s[a][b] == 1;
      }
    if (!any1) return 2;
    const float k = d.resp[p];
    bool wait = false;
#pragma unroll
    for (int a = 0; a < 7; ++a)
#pragma unroll
      for (int b = 0; b < 7; ++b)
        if (s[a][b] == 1 && outranks(d.resp[size_t(y + a - 3) * W + (x + b - 3)], k, a - 3, b - 3))
          wait = true;
    return wait ? 1 : 2;
  }
  const int y = int(p / uint32_t(W)), x = int(p - uint32_t(y) * uint32_t(W));
  const float k = d.resp[p];
  bool wait = false;
  for (int dy = -rad; dy <= rad; ++dy) {
    const int py = y + dy;
    if (py < 0 || py >= H) continue;
    const volatile uint8_t* srow = d.state + size_t(py) * W;
    for (int dx = -rad; dx <= rad; ++dx) {
      const int px = x + dx;
      if (px < 0 || px >= W || (dx == 0 && dy == 0)) continue;
      const uint8_t s = srow[px];
      if (s == 2) return 0;
      if (s == 1 && !wait && outranks(d.resp[size_t(py) * W + px], k, dy, dx)) wait = true;
    }
  }
  return wait ? 1 : 2;
}

// Append n pixel indices held in LDS to the frame's kept list (one atomic).
__device__ __forceinline__ void append_kept(const DetDesc& d, const uint32_t* newk, int n,
                                            int* base_sh) {
  if (threadIdx.x == 0 && n > 0) *base_sh = atomicAdd(&d.cnt[1], n);
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x) d.kept[*base_sh + j] = newk[j];
}

// One suppression pass over a 64x16 tile: the tile's undecided candidates and a
// halo of `rad` pixels are staged in LDS. Only neighbours that outrank a
// candidate can decide it (a kept neighbour always outranks: it was decided
// while this candidate was undecided). For radii up to 3 the candidates are
// ranked by priority and swept in that order, 64 at a time: each lane reads its
// window once (in-tile neighbours compare by rank, halo ones by response), and
// ballot rounds resolve the chunk (kept / removed / stuck behind an undecided
// neighbour outside the tile), which reaches the fixed point of the parallel
// rule in one sweep. Wider radii iterate the rule in LDS instead. Decided
// states go back to global memory, newly kept candidates to the kept list;
// pass 0 counts the candidates, the last pass hands the undecided ones to the
// finisher.
__global__ __launch_bounds__(kDetThreads) void k_det_nms_tile(const DetDesc* __restrict__ frames,
                                                              int* __restrict__ flags, int pass,
                                                              int last) {
  if (flags[pass] == 0) return;
  const DetDesc d = frames[blockIdx.z];
  const int W = d.W, H = d.H, rad = d.rad;
  const int x0 = blockIdx.x * kNtW, y0 = blockIdx.y * kNtH;
  // uniform over the block; the host launches this kernel only for rad <= kNtMaxRad
  if (x0 >= W || y0 >= H || rad > kNtMaxRad) return;
  // tiles that ended the previous pass with nothing undecided stay done
  const int tile = blockIdx.y * ((W + kNtW - 1) / kNtW) + blockIdx.x;
  const int nt = ((W + kNtW - 1) / kNtW) * ((H + kNtH - 1) / kNtH);
  if (pass > 0 && d.tpend[((pass - 1) & 1) * nt + tile] == 0) {
    if (threadIdx.x == 0) d.tpend[(pass & 1) * nt + tile] = 0;
    return;
  }
  constexpr int kN = kNtW * kNtH;
  __shared__ uint8_t st[kNtLH * kNtLW];
  __shared__ float rs[kNtLH * kNtLW];
  __shared__ uint16_t rank[kNtLH * kNtLW];
  __shared__ uint16_t list[kN];
  __shared__ uint64_t key[kN];
  __shared__ uint16_t ord[kN];
  __shared__ uint32_t newk[kN];
  __shared__ int nlist, nnew, npend, changed, kbase;
  const int t = threadIdx.x;
  if (t == 0) {
    nlist = 0;
    nnew = 0;
    npend = 0;
  }
  const int LW = kNtW + 2 * rad, LH = kNtH + 2 * rad;
  const bool sweep = rad <= 3;
  __syncthreads();
  // stage the region: all state loads in flight together, then the responses
  constexpr int kPer = (kNtLW * kNtLH + kDetThreads - 1) / kDetThreads;
  uint8_t sv[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int i = t + u * kDetThreads;
    const int ly = i / LW, lx = i - ly * LW;
    const int gx = x0 - rad + lx, gy = y0 - rad + ly;
    sv[u] = (i < LW * LH && gx >= 0 && gx < W && gy >= 0 && gy < H)
                ? d.state[size_t(gy) * W + gx] : uint8_t(0);
  }
  float rv[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int i = t + u * kDetThreads;
    const int ly = i / LW, lx = i - ly * LW;
    rv[u] = sv[u] == 1 ? d.resp[size_t(y0 - rad + ly) * W + (x0 - rad + lx)] : 0.f;
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int i = t + u * kDetThreads;
    if (i >= LW * LH) break;
    const int ly = i / LW, lx = i - ly * LW;
    st[i] = sv[u];
    rs[i] = rv[u];
    rank[i] = 0xFFFF;
    if (sv[u] == 1 && lx >= rad && lx < rad + kNtW && ly >= rad && ly < rad + kNtH)
      list[atomicAdd(&nlist, 1)] = uint16_t(i);
  }
  __syncthreads();
  const int n = nlist;
  if (pass == 0 && t == 0 && n > 0) atomicAdd(&d.cnt[0], n);
  auto global_of = [&](int i) {
    const int ly = i / LW, lx = i - ly * LW;
    return uint32_t((y0 - rad + ly) * W + (x0 - rad + lx));
  };
  if (sweep && n > 0) {
    for (int j = t; j < n; j += kDetThreads) {
      // priority key: response, then earlier scan position inside the tile
      // (row-major, as the image's), then j in the low 10 bits
      const int i = list[j];
      const int ly = i / LW, lx = i - ly * LW;
      const int pos = (ly - rad) * kNtW + (lx - rad);
      key[j] = (uint64_t(ordered_f32(rs[i])) << 32) | (uint64_t(kN - 1 - pos) << 10) |
               uint64_t(j);
    }
    __syncthreads();
    // rank by counting larger keys (keys are distinct), then invert
    for (int j = t; j < n; j += kDetThreads) {
      const uint64_t kj = key[j];
      int r = 0, q = 0;
      for (; q + 8 <= n; q += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) r += key[q + u] > kj ? 1 : 0;
      }
      for (; q < n; ++q) r += key[q] > kj ? 1 : 0;
      rank[list[j]] = uint16_t(r);
      ord[r] = uint16_t(j);
    }
    __syncthreads();
    // sweep in rank order, 64 candidates at a time; the window of each is read
    // once: kept neighbours remove it, in-tile neighbours of higher rank in this
    // chunk become its lane mask, undecided ones of earlier chunks or of the halo
    // (compared by response) leave it stuck
    const int wv = t / kWave, lane = __lane_id();
    for (int c0 = 0; c0 < n; c0 += kWave) {
      if (wv == 0) {
        const int k = c0 + lane;
        const bool act = k < n;
        uint64_t nb = 0;
        bool rf = false, sf = false;
        int i = 0;
        if (act) {
          i = list[ord[k]];
          const float rk = rs[i];
          for (int dy = -rad; dy <= rad; ++dy) {
            const int row = i + dy * LW;
#pragma unroll
            for (int b = 0; b < 7; ++b) {
              const int dx = b - 3;
              const bool in = abs(dx) <= rad && (dx | dy) != 0;
              const int q = in ? row + dx : i;
              const uint8_t sq = st[q];
              const int rq = rank[q];
              const float fq = rs[q];
              if (!in || sq == 0) continue;
              if (sq == 2) {
                rf = true;
              } else if (rq != 0xFFFF) {
                if (rq < k) {
                  if (rq >= c0) nb |= 1ull << (rq - c0);
                  else sf = true;
                }
              } else if (outranks(fq, rk, dy, dx)) {
                sf = true;
              }
            }
          }
        }
        // resolve the chunk: a lane decides once its in-chunk outranking
        // neighbours have; the highest undecided lane always can
        uint64_t und = __ballot(act), kept = 0, stuck = 0;
        while (und) {
          const bool mine = act && ((und >> lane) & 1ull) && (nb & und) == 0;
          const bool rm = rf || (nb & kept) != 0;
          const bool sk = !rm && (sf || (nb & stuck) != 0);
          const uint64_t dec = __ballot(mine);
          kept |= __ballot(mine && !rm && !sk);
          stuck |= __ballot(mine && sk);
          und &= ~dec;
        }
        if (act) {
          const bool kp = (kept >> lane) & 1ull, sk = (stuck >> lane) & 1ull;
          st[i] = kp ? 2 : (sk ? 1 : 0);
          if (kp) newk[atomicAdd(&nnew, 1)] = global_of(i);
        }
      }
      __syncthreads();
    }
  } else if (n > 0) {
    for (;;) {
      if (t == 0) changed = 0;
      __syncthreads();
      for (int j = t; j < n; j += kDetThreads) {
        const int i = list[j];
        if (st[i] != 1) continue;
        int dec = 2;
        const float k = rs[i];
        const int ly = i / LW, lx = i - ly * LW;
        for (int dy = -rad; dy <= rad && dec != 0; ++dy) {
          const volatile uint8_t* srow = st + (ly + dy) * LW + lx;
          const float* rrow = rs + (ly + dy) * LW + lx;
          for (int dx = -rad; dx <= rad; ++dx) {
            if (dx == 0 && dy == 0) continue;
            const uint8_t s = srow[dx];
            if (s == 2) {
              dec = 0;
              break;
            }
            if (s == 1 && dec == 2 && outranks(rrow[dx], k, dy, dx)) dec = 1;
          }
        }
        if (dec != 1) {
          st[i] = uint8_t(dec);
          changed = 1;
          if (dec == 2) newk[atomicAdd(&nnew, 1)] = global_of(i);
        }
      }
      __syncthreads();
      const int c = changed;
      __syncthreads();
      if (!c) break;
    }
  }
  bool pend = false;
  for (int j = t; j < n; j += kDetThreads) {
    const int i = list[j];
    const uint8_t s = st[i];
    const uint32_t g = global_of(i);
    if (s == 1) {
      pend = true;
      if (last) newk[kN - 1 - atomicAdd(&npend, 1)] = g;   // from the top
    } else {
      d.state[g] = s;
    }
  }
  if (pend) flags[pass + 1] = 1;
  const int any_pend = __syncthreads_or(pend);
  if (t == 0) d.tpend[(pass & 1) * nt + tile] = any_pend;
  append_kept(d, newk, nnew, &kbase);
  if (last) {
    __syncthreads();
    const int np = npend;
    if (t == 0 && np > 0) kbase = atomicAdd(&d.cnt[3], np);
    __syncthreads();
    for (int j = t; j < np; j += kDetThreads) d.cand[kbase + j] = newk[kN - 1 - j];
  }
}

// One suppression round over the candidate list from global memory (radii too
// wide for the LDS tile).
__global__ __launch_bounds__(kDetThreads) void k_det_nms(const DetDesc* __restrict__ frames,
                                                         int* __restrict__ flags, int round) {
  if (flags[round] == 0) return;
  const DetDesc d = frames[blockIdx.y];
  const int n = d.cnt[0];
  bool pending = false;
  for (int i = blockIdx.x * kDetThreads + threadIdx.x; i < n; i += gridDim.x * kDetThreads) {
    const uint32_t p = d.cand[i];
    if (d.state[p] != 1) continue;
    const int s = nms_decide(d, p);
    if (s == 1) {
      pending = true;
    } else {
      d.state[p] = uint8_t(s);
      if (s == 2) d.kept[atomicAdd(&d.cnt[1], 1)] = p;
    }
  }
  if (pending) flags[round + 1] = 1;
}

// Passes inside one workgroup per frame until every candidate is decided. Each
// pass decides at least the undecided candidate of highest priority. The
// undecided candidates are gathered into LDS first (kFinList of them; beyond
// that the passes walk the whole candidate list).
__global__ __launch_bounds__(kFinThreads) void k_det_nms_fin(const DetDesc* __restrict__ frames,
                                                             const int* __restrict__ flags,
                                                             int round) {
  if (flags[round] == 0) return;
  const DetDesc d = frames[blockIdx.x];
  const int n = d.tiled ? d.cnt[3] : d.cnt[0];
  __shared__ uint32_t open[kFinList];
  __shared__ int n_open, pending;
  if (threadIdx.x == 0) n_open = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kFinThreads) {
    const uint32_t p = d.cand[i];
    if (d.state[p] == 1) {
      const int k = atomicAdd(&n_open, 1);
      if (k < kFinList) open[k] = p;
    }
  }
  __syncthreads();
  const bool in_lds = n_open <= kFinList;
  const int m = in_lds ? n_open : n;
  for (;;) {
    if (threadIdx.x == 0) pending = 0;
    __syncthreads();
    bool mine = false;
    for (int i = threadIdx.x; i < m; i += kFinThreads) {
      const uint32_t p = in_lds ? open[i] : d.cand[i];
      if (d.state[p] != 1) continue;
      const int s = nms_decide(d, p);
      if (s == 1) {
        mine = true;
      } else {
        d.state[p] = uint8_t(s);
        if (s == 2) d.kept[atomicAdd(&d.cnt[1], 1)] = p;
      }
    }
    if (mine) pending = 1;
    __threadfence();   // this pass's decisions reach L2 before the next pass reads them
    __syncthreads();
    const int again = pending;
    __syncthreads();
    if (!again) break;
  }
}

// Texture floor and sub-pixel refinement of a kept candidate at (x, y) with
// grad_window radius 1, from its 7x7 neighbourhood held in registers (inside
// the image: candidates lie at least 5 pixels from every edge). The 7x7 sums
// are integers, so int arithmetic gives the reference's double sums exactly.
__device__ __forceinline__ float refine_r1(const uint8_t* img, int W, int x, int y, float c,
                                           int subpixel, float min_std, float* kx, float* ky) {
  int v[7][7];
#pragma unroll
  for (int a = 0; a < 7; ++a)
#pragma unroll
    for (int b = 0; b < 7; ++b) v[a][b] = img[size_t(y - 3 + a) * W + (x - 3 + b)];
  int s1 = 0, s2 = 0;
#pragma unroll
  for (int a = 0; a < 7; ++a)
#pragma unroll
    for (int b = 0; b < 7; ++b) {
      s1 += v[a][b];
      s2 += v[a][b] * v[a][b];
    }
  const double mean = double(s1) / 49.0;
  const double var = double(s2) / 49.0 - mean * mean;
  const float ls = var > 0.0 ? float(sqrt(var)) : 0.f;
  if (ls < min_std || !subpixel) return ls;
  // shi_tomasi_at(x + ox, y + oy, 3) from the registers
  auto shi = [&](int ox, int oy) {
    auto px = [&](int dx, int dy) { return v[3 + oy + dy][3 + ox + dx]; };
    float sxx = 0.f, syy = 0.f, sxy = 0.f;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const float gx = (float(px(dx + 1, dy - 1)) - float(px(dx - 1, dy - 1))
                          + 2.f * (float(px(dx + 1, dy)) - float(px(dx - 1, dy)))
                          + float(px(dx + 1, dy + 1)) - float(px(dx - 1, dy + 1))) * 0.125f;
        const float gy = (float(px(dx - 1, dy + 1)) - float(px(dx - 1, dy - 1))
                          + 2.f * (float(px(dx, dy + 1)) - float(px(dx, dy - 1)))
                          + float(px(dx + 1, dy + 1)) - float(px(dx + 1, dy - 1))) * 0.125f;
        sxx += gx * gx;
        syy += gy * gy;
        sxy += gx * gy;
      }
    const float n = 1.f / 9.f;
    const float a = sxx * n, b = syy * n, cc = sxy * n;
    const float half_tr = 0.5f * (a + b);
    const float q = 0.25f * (a - b) * (a - b) + cc * cc;
    return half_tr - sqrtf((0.f < q) ? q : 0.f);
  };
  const float l = shi(-1, 0), rr = shi(1, 0);
  const float dxx = l - 2.f * c + rr;
  if (fabsf(dxx) > 1e-12f) {
    const float o = 0.5f * (l - rr) / dxx;
    if (fabsf(o) <= 0.5f) *kx += o;
  }
  const float u = shi(0, -1), dn = shi(0, 1);
  const float dyy = u - 2.f * c + dn;
  if (fabsf(dyy) > 1e-12f) {
    const float o = 0.5f * (u - dn) / dyy;
    if (fabsf(o) <= 0.5f) *ky += o;
  }
  return ls;
}

// Texture floor, sub-pixel parabola and cell count of each kept candidate.
__global__ __launch_bounds__(kDetThreads) void k_det_refine(const DetDesc* __restrict__ frames) {
  const DetDesc d = frames[blockIdx.y];
  const int n = d.cnt[1];
  const int W = d.W, H = d.H, r = d.r;
  for (int j = blockIdx.x * kDetThreads + threadIdx.x; j < n; j += gridDim.x * kDetThreads) {
    const uint32_t p = d.kept[j];
    const int y = int(p / uint32_t(W)), x = int(p - uint32_t(y) * uint32_t(W));
    const float c = d.resp[p];
    float kx = float(x), ky = float(y), ls;
    if (r == 1 && x >= 3 && y >= 3 && x + 3 < W && y + 3 < H) {
      ls = refine_r1(d.img, W, x, y, c, d.subpixel, d.min_std, &kx, &ky);
    } else {
      ls = local_std7(d.img, W, H, x, y);
      if (!(ls < d.min_std) && d.subpixel) {
        const float l = shi_global(d.img, W, H, x - 1, y, r);
        const float rr = shi_global(d.img, W, H, x + 1, y, r);
        const float dxx = l - 2.f * c + rr;
        if (fabsf(dxx) > 1e-12f) {
          const float o = 0.5f * (l - rr) / dxx;
          if (fabsf(o) <= 0.5f) kx += o;
        }
        const float u = shi_global(d.img, W, H, x, y - 1, r);
        const float dn = shi_global(d.img, W, H, x, y + 1, r);
        const float dyy = u - 2.f * c + dn;
        if (fabsf(dyy) > 1e-12f) {
          const float o = 0.5f * (u - dn) / dyy;
          if (fabsf(o) <= 0.5f) ky += o;
        }
      }
    }
    int s = -1;
    if (!(ls < d.min_std)) {
      s = atomicAdd(&d.cell_cnt[(y / d.cell) * d.cols + (x / d.cell)], 1);
      d.rec[j] = make_float4(kx, ky, c, ls);
    }
    d.slot[j] = s;
  }
}

// Exclusive scans of the cell counts (cell lists) and of min(count, keep)
// (output places); one workgroup per frame.
__global__ __launch_bounds__(kFinThreads) void k_det_scan(const DetDesc* __restrict__ frames) {
  const DetDesc d = frames[blockIdx.x];
  const int nc = d.ncells, keep = d.keep;
  const int per = (nc + kFinThreads - 1) / kFinThreads;
  const int c0 = min(nc, int(threadIdx.x) * per), c1 = min(nc, c0 + per);
  int a = 0, o = 0;
  for (int c = c0; c < c1; ++c) {
    const int v = d.cell_cnt[c];
    a += v;
    o += min(v, keep);
  }
  const int t = int(threadIdx.x);
  __shared__ int sa[kFinThreads], so[kFinThreads];
  sa[t] = a;
  so[t] = o;
  __syncthreads();
  for (int off = 1; off < kFinThreads; off <<= 1) {
    const int va = t >= off ? sa[t - off] : 0;
    const int vo = t >= off ? so[t - off] : 0;
    __syncthreads();
    sa[t] += va;
    so[t] += vo;
    __syncthreads();
  }
  int ra = sa[t] - a, ro = so[t] - o;
  for (int c = c0; c < c1; ++c) {
    const int v = d.cell_cnt[c];
    d.cell_off[c] = ra;
    d.out_off[c] = ro;
    ra += v;
    ro += min(v, keep);
  }
  if (t == kFinThreads - 1) {
    d.cell_off[nc] = sa[kFinThreads - 1];
    d.out_off[nc] = so[kFinThreads - 1];
    d.cnt[2] = so[kFinThreads - 1];
    *d.n_out = so[kFinThreads - 1];
  }
}

__global__ __launch_bounds__(kDetThreads) void k_det_scatter(const DetDesc* __restrict__ frames) {
  const DetDesc d = frames[blockIdx.y];
  const int n = d.cnt[1];
  for (int j = blockIdx.x * kDetThreads + threadIdx.x; j < n; j += gridDim.x * kDetThreads) {
    const int s = d.slot[j];
    if (s < 0) continue;
    const uint32_t p = d.kept[j];
    const int y = int(p / uint32_t(d.W)), x = int(p - uint32_t(y) * uint32_t(d.W));
    d.items[d.cell_off[(y / d.cell) * d.cols + (x / d.cell)] + s] = j;
  }
}

// One wave per cell: rank each member by (response desc, pixel index asc) and
// write the first `keep` ranks to the output in rank order.
__global__ __launch_bounds__(kDetThreads) void k_det_select(const DetDesc* __restrict__ frames) {
  const DetDesc d = frames[blockIdx.y];
  const int cell = blockIdx.x * (kDetThreads / kWave) + int(threadIdx.x / kWave);
  if (cell >= d.ncells) return;   // uniform over the wave
  const int lane = __lane_id();
  const int off = d.cell_off[cell], m = d.cell_off[cell + 1] - off;
  const int base = d.out_off[cell], keep = d.keep;
  for (int jb = 0; jb < m; jb += kWave) {
    const int j = jb + lane;
    const bool mine = j < m;
    int ij = 0;
    float rj = 0.f;
    uint32_t pj = 0;
    if (mine) {
      ij = d.items[off + j];
      rj = d.rec[ij].z;
      pj = d.kept[ij];
    }
    int rank = 0;
    for (int tb = 0; tb < m; tb += kWave) {
      const int t = tb + lane;
      float rt = 0.f;
      uint32_t pt = 0;
      if (t < m) {
        const int it = d.items[off + t];
        rt = d.rec[it].z;
        pt = d.kept[it];
      }
      const int lim = min(kWave, m - tb);
      for (int u = 0; u < lim; ++u) {
        const float ru = __shfl(rt, u);
        const uint32_t pu = __shfl(pt, u);
        rank += (ru > rj || (ru == rj && pu < pj)) ? 1 : 0;
      }
    }
    if (mine && rank < keep) d.out[base + rank] = d.rec[ij];
  }
}

// census_at (preproc.cpp:142-162) at round-to-nearest keypoint positions,
// int(x + 0.5f) and int(y + 0.5f) as de_pipe.cpp:168-172 computes them.
__global__ void k_census_kp(const uint8_t* __restrict__ img, int W, int H, int hw, int hh,
                            const float4* __restrict__ kp, const int* __restrict__ n_kp,
                            int n_host, uint64_t* __restrict__ desc) {
  const int n = n_kp ? *n_kp : n_host;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const float4 q = kp[k];
    const int x = int(q.x + 0.5f), y = int(q.y + 0.5f);
    uint64_t bits = 0;
    const bool inside = x >= hw && y >= hh && x < W - hw && y < H - hh;
    if (inside && (2 * hw + 1) * (2 * hh + 1) - 1 <= 64) {
      const uint8_t c = img[size_t(y) * W + x];
      int bit = 0;
      for (int dy = -hh; dy <= hh; ++dy) {
        const uint8_t* row = img + size_t(y + dy) * W + x;
        for (int dx = -hw; dx <= hw; ++dx) {
          if (dx == 0 && dy == 0) continue;
          if (row[dx] < c) bits |= 1ull << bit;
          ++bit;
        }
      }
    }
    desc[k] = bits;
  }
}

// SAD over the (2h+1)^2 windows at (xl, y) and (xl - d, y); -1 near a border.
__device__ long sad_cost(const uint8_t* L, const uint8_t* R, int W, int H, int xl, int y, int d,
                         int h) {
  const int xr = xl - d;
  if (xl - h < 0 || xl + h >= W) return -1;
  if (xr - h < 0 || xr + h >= W) return -1;
  if (y - h < 0 || y + h >= H) return -1;
  long acc = 0;
  for (int v = -h; v <= h; ++v) {
    const uint8_t* lr = L + size_t(y + v) * W;
    const uint8_t* rr = R + size_t(y + v) * W;
    for (int u = -h; u <= h; ++u) acc += abs(int(lr[xl + u]) - int(rr[xr + u]));
  }
  return acc;
}

// refine_disparity (match.cpp:768-810) for one match: parabola through the SAD
// costs at round(d) - 1, round(d), round(d) + 1, vertex clamped to +-0.5.
__device__ void refine_one(const uint8_t* L, const uint8_t* R, int W, int H, const float4* kl,
                           int nl, int nr, int h, de_match* m) {
  if (m->left < 0 || m->right < 0 || m->left >= nl || m->right >= nr) return;
  const float4 q = kl[m->left];
  const int xl = int(q.x + 0.5f), y = int(q.y + 0.5f);
  const int d0 = int(roundf(m->disparity));
  const long c0 = sad_cost(L, R, W, H, xl, y, d0, h);
  const long cm = sad_cost(L, R, W, H, xl, y, d0 - 1, h);
  const long cp = sad_cost(L, R, W, H, xl, y, d0 + 1, h);
  if (c0 < 0 || cm < 0 || cp < 0) return;
  const double denom = double(cm) - 2.0 * double(c0) + double(cp);
  if (denom <= 0.0) return;
  double off = 0.5 * (double(cm) - double(cp)) / denom;
  if (off > 0.5) off = 0.5;
  if (off < -0.5) off = -0.5;
  m->disparity = float(double(d0) + off);
}

__global__ void k_refine(const uint8_t* __restrict__ L, const uint8_t* __restrict__ R, int W,
                         int H, const float4* __restrict__ kl, int nl, int nr, int h,
                         de_match* __restrict__ m, int n) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    refine_one(L, R, W, H, kl, nl, nr, h, &m[k]);
}

// de_pipe's tail (de_pipe.cpp:175-199) in one workgroup: drop matches whose
// margin is below min_margin (order kept), refine the rest when asked, and
// write the matches and the DEMR records (x_left, y_left, disparity, margin).
__global__ __launch_bounds__(kFinThreads) void k_pipe_finish(
    const uint8_t* __restrict__ L, const uint8_t* __restrict__ R, int W, int H,
    const float4* __restrict__ kl, int nl, int nr, const de_match* __restrict__ in,
    const int* __restrict__ n_in, float min_margin, int subpixel, int h,
    de_match* __restrict__ out, float4* __restrict__ rec, int* __restrict__ n_out) {
  const int n = *n_in;
  __shared__ int ws[kFinThreads / kWave];
  __shared__ int total;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  for (int b0 = 0; b0 < n; b0 += kFinThreads) {
    const int k = b0 + threadIdx.x;
    de_match m;
    bool keep = false;
    if (k < n) {
      m = in[k];
      keep = !(min_margin > 0.f && m.margin < min_margin);
    }
    const uint64_t bal = __ballot(keep);
    if (lane == 0) ws[wid] = __popcll(bal);
    __syncthreads();
    int before = total;
    for (int w = 0; w < wid; ++w) before += ws[w];
    const int pos = before + __popcll(bal & lanes_below());
    if (keep) {
      if (subpixel) refine_one(L, R, W, H, kl, nl, nr, h, &m);
      out[pos] = m;
      const float4 q = (m.left >= 0 && m.left < nl) ? kl[m.left] : make_float4(0.f, 0.f, 0.f, 0.f);
      rec[pos] = make_float4(q.x, q.y, m.disparity, m.margin);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int s = 0;
      for (int w = 0; w < kFinThreads / kWave; ++w) s += ws[w];
      total += s;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *n_out = total;
}

__global__ void k_fast_map(const uint8_t* __restrict__ img, int W, int H, int thr, int r,
                           uint8_t* __restrict__ corner, float* __restrict__ resp) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= W) return;
  const bool inside = x >= 3 && y >= 3 && x < W - 3 && y < H - 3;
  corner[size_t(y) * W + x] = (inside && fast_at(GlobalPx{img, W, x, y}, thr)) ? 1 : 0;
  resp[size_t(y) * W + x] = shi_global(img, W, H, x, y, r);
}

size_t a256(size_t v) { return (v + 255) & ~size_t(255); }

// Geometry of one frame from its configuration.
struct DetGeom {
  int border, r, rad, cell, cols, rows, keep, ncells;
  size_t region;   // pixels FAST examines: an upper bound on candidates
  size_t max_out;  // upper bound on keypoints
};

DetGeom det_geom(const de_detector_cfg& c, int W, int H) {
  DetGeom g;
  g.border = std::max(c.border, 5);
  g.r = std::max(1, c.grad_window / 2);
  g.rad = std::max(1, c.nms_radius);
  g.cell = std::max(4, c.cell);
  g.cols = (W + g.cell - 1) / g.cell;
  g.rows = (H + g.cell - 1) / g.cell;
  g.keep = std::max(1, c.per_cell);
  g.ncells = g.cols * g.rows;
  g.region = size_t(std::max(0, W - 2 * g.border)) * size_t(std::max(0, H - 2 * g.border));
  g.max_out = std::min(size_t(g.ncells) * size_t(g.keep), g.region);
  return g;
}

size_t det_tiles(int W, int H) {
  return size_t((W + kNtW - 1) / kNtW) * size_t((H + kNtH - 1) / kNtH);
}

size_t det_frame_bytes(int W, int H, const DetGeom& g) {
  const size_t px = size_t(W) * size_t(H), nb = std::max<size_t>(g.region, 1);
  const size_t nc = size_t(g.ncells) + 1;
  return a256(px) + a256(px * 4) + a256(nb * 4) * 2 + a256(nb * 16) + a256(nb * 4) +
         a256(nb * 4) + a256(nc * 4) * 3 + a256(16) + a256(det_tiles(W, H) * 8);
}

void det_carve(void* base, int W, int H, const DetGeom& g, DetDesc* d) {
  uint8_t* p = static_cast<uint8_t*>(base);
  auto take = [&](size_t bytes) { void* r = p; p += a256(bytes); return r; };
  const size_t px = size_t(W) * size_t(H), nb = std::max<size_t>(g.region, 1);
  const size_t nc = size_t(g.ncells) + 1;
  d->state = static_cast<uint8_t*>(take(px));
  d->resp = static_cast<float*>(take(px * 4));
  d->cand = static_cast<uint32_t*>(take(nb * 4));
  d->kept = static_cast<uint32_t*>(take(nb * 4));
  d->rec = static_cast<float4*>(take(nb * 16));
  d->slot = static_cast<int*>(take(nb * 4));
  d->items = static_cast<int*>(take(nb * 4));
  d->cell_cnt = static_cast<int*>(take(nc * 4));
  d->cell_off = static_cast<int*>(take(nc * 4));
  d->out_off = static_cast<int*>(take(nc * 4));
  d->cnt = static_cast<int*>(take(16));
  d->tpend = static_cast<int*>(take(det_tiles(W, H) * 8));
}

void det_fill(const de_detector_cfg& c, const DetGeom& g, const uint8_t* img, int W, int H,
              DetDesc* d) {
  d->img = img;
  d->W = W;
  d->H = H;
  d->border = g.border;
  d->thr = c.fast_threshold;
  d->r = g.r;
  d->rad = g.rad;
  d->cell = g.cell;
  d->cols = g.cols;
  d->rows = g.rows;
  d->keep = g.keep;
  d->subpixel = c.subpixel ? 1 : 0;
  d->ncells = g.ncells;
  d->min_resp = c.min_response;
  d->min_std = c.min_local_std;
}

}  // namespace

int det_validate(const de_detector_cfg& c, int W, int H) {
  if (W < 0 || H < 0 || (W > 0 && H > 0 && size_t(W) * size_t(H) > (size_t(1) << 31)))
    return fail(DE_ERR_ARG, "detector: image %dx%d", W, H);
  if (std::max(1, c.nms_radius) > kMaxNmsRadius)
    return fail(DE_ERR_UNSUPPORTED, "detector: nms_radius %d exceeds %d", c.nms_radius,
                kMaxNmsRadius);
  if (c.grad_window / 2 > 4096) return fail(DE_ERR_ARG, "detector: grad_window %d", c.grad_window);
  return DE_OK;
}

size_t det_max_keypoints(const de_detector_cfg& c, int W, int H) {
  return det_geom(c, W, H).max_out;
}

// Launch the detector over n frames whose scratch the context's det buffer
// holds; outputs (device) in frames[k].out / n_out. No synchronisation.
int det_launch(Ctx& x, const de_detect_frame* frames, const de_detector_cfg* cfgs, int n,
               hipStream_t s, hipEvent_t* ev) {
  std::vector<DetDesc> fr(static_cast<size_t>(n));
  std::vector<DetGeom> geo(static_cast<size_t>(n));
  size_t total = a256(kFlags * sizeof(int));
  for (int k = 0; k < n; ++k) {
    geo[size_t(k)] = det_geom(cfgs[k], frames[k].W, frames[k].H);
    total += a256(det_frame_bytes(frames[k].W, frames[k].H, geo[size_t(k)]));
  }
  if (int rc = x.det_scratch.ensure(total)) return rc;
  uint8_t* p = static_cast<uint8_t*>(x.det_scratch.ptr);
  int* flags = reinterpret_cast<int*>(p);
  p += a256(kFlags * sizeof(int));
  int maxW = 1, maxH = 1, max_cells = 1;
  size_t max_region = 1;
  bool tiled = true;   // every frame's radius fits the LDS suppression tile
  for (int k = 0; k < n; ++k) {
    const de_detect_frame& q = frames[k];
    DetDesc& d = fr[size_t(k)];
    std::memset(&d, 0, sizeof(d));
    det_fill(cfgs[k], geo[size_t(k)], q.img, q.W, q.H, &d);
    det_carve(p, q.W, q.H, geo[size_t(k)], &d);
    p += a256(det_frame_bytes(q.W, q.H, geo[size_t(k)]));
    d.out = reinterpret_cast<float4*>(q.out);
    d.n_out = q.n_out;
    maxW = std::max(maxW, q.W);
    maxH = std::max(maxH, q.H);
    max_cells = std::max(max_cells, geo[size_t(k)].ncells);
    max_region = std::max(max_region, geo[size_t(k)].region);
    tiled = tiled && geo[size_t(k)].rad <= kNtMaxRad;
  }
  for (DetDesc& d : fr) d.tiled = tiled ? 1 : 0;
  Ctx::Slot* slot;
  const size_t db = fr.size() * sizeof(DetDesc);
  if (int rc = x.acquire_slot(db, &slot)) return rc;
  std::memcpy(slot->host, fr.data(), db);
  DE_HIP(hipMemcpyAsync(slot->dev, slot->host, db, hipMemcpyHostToDevice, s));
  const DetDesc* dd = static_cast<const DetDesc*>(slot->dev);
  const int cu = 256;
  const int list_blocks = int(std::min<size_t>((max_region + kDetThreads - 1) / kDetThreads,
                                               size_t(4 * cu)));
  const int cnt_blocks = std::max(1, std::min(list_blocks, (4 * cu + n - 1) / n));
  if (ev) DE_HIP(hipEventRecord(ev[0], s));
  if (!tiled)
    hipLaunchKernelGGL(k_det_init,
                       dim3((std::max(max_cells, kFlags) + kDetThreads - 1) / kDetThreads, n),
                       dim3(kDetThreads), 0, s, dd, flags);
  hipLaunchKernelGGL(k_det_fast, dim3((maxW + kTileW - 1) / kTileW, (maxH + kTileH - 1) / kTileH, n),
                     dim3(kDetThreads), 0, s, dd, flags);
  DE_HIP(hipGetLastError());
  if (ev) DE_HIP(hipEventRecord(ev[1], s));
  const int passes = tiled ? kNmsTilePasses : kNmsRounds;
  for (int r = 0; r < passes; ++r) {
    if (tiled)
      hipLaunchKernelGGL(k_det_nms_tile, dim3((maxW + kNtW - 1) / kNtW, (maxH + kNtH - 1) / kNtH, n),
                         dim3(kDetThreads), 0, s, dd, flags, r, r == passes - 1 ? 1 : 0);
    else
      hipLaunchKernelGGL(k_det_nms, dim3(cnt_blocks, n), dim3(kDetThreads), 0, s, dd, flags, r);
  }
  hipLaunchKernelGGL(k_det_nms_fin, dim3(n), dim3(kFinThreads), 0, s, dd, flags, passes);
  DE_HIP(hipGetLastError());
  if (ev) DE_HIP(hipEventRecord(ev[2], s));
  hipLaunchKernelGGL(k_det_refine, dim3(cnt_blocks, n), dim3(kDetThreads), 0, s, dd);
  hipLaunchKernelGGL(k_det_scan, dim3(n), dim3(kFinThreads), 0, s, dd);
  hipLaunchKernelGGL(k_det_scatter, dim3(cnt_blocks, n), dim3(kDetThreads), 0, s, dd);
  hipLaunchKernelGGL(k_det_select, dim3((max_cells + 3) / 4, n), dim3(kDetThreads), 0, s, dd);
  DE_HIP(hipGetLastError());
  if (ev) DE_HIP(hipEventRecord(ev[3], s));
  DE_HIP(hipEventRecord(slot->done, s));
  slot->used = true;
  return DE_OK;
}

// Device counters of frame k of the last det_launch: candidates, kept.
static int det_counts(Ctx& x, const de_detect_frame* frames, const de_detector_cfg* cfgs, int k,
                      int* counts) {
  size_t off = a256(kFlags * sizeof(int));
  for (int j = 0; j < k; ++j)
    off += a256(det_frame_bytes(frames[j].W, frames[j].H, det_geom(cfgs[j], frames[j].W, frames[j].H)));
  DetDesc d;
  const DetGeom g = det_geom(cfgs[k], frames[k].W, frames[k].H);
  det_carve(static_cast<uint8_t*>(x.det_scratch.ptr) + off, frames[k].W, frames[k].H, g, &d);
  DE_HIP(hipMemcpyAsync(counts, d.cnt, 2 * sizeof(int), hipMemcpyDeviceToHost, x.stream));
  return DE_OK;
}

// Stage the images of n frames (host pointers) in the det io buffer, run the
// detector, return device output pointers (inside the same buffer).
static int det_host_run(Ctx& x, const uint8_t* const* imgs, const int* Ws, const int* Hs,
                        const de_detector_cfg* cfgs, int n, std::vector<de_detect_frame>* fr,
                        hipEvent_t* ev) {
  size_t bytes = 0;
  for (int k = 0; k < n; ++k)
    bytes += a256(size_t(Ws[k]) * size_t(Hs[k])) +
             a256(std::max<size_t>(det_max_keypoints(cfgs[k], Ws[k], Hs[k]), 1) * 16) + 256;
  if (int rc = x.det_io.ensure(bytes)) return rc;
  uint8_t* p = static_cast<uint8_t*>(x.det_io.ptr);
  fr->assign(static_cast<size_t>(n), de_detect_frame{});
  for (int k = 0; k < n; ++k) {
    de_detect_frame& q = (*fr)[size_t(k)];
    const size_t px = size_t(Ws[k]) * size_t(Hs[k]);
    q.img = p;
    if (px) DE_HIP(hipMemcpyAsync(p, imgs[k], px, hipMemcpyHostToDevice, x.stream));
    p += a256(px);
    q.W = Ws[k];
    q.H = Hs[k];
    q.out = reinterpret_cast<de_keypoint*>(p);
    q.cap = int(std::max<size_t>(det_max_keypoints(cfgs[k], Ws[k], Hs[k]), 1));
    p += a256(size_t(q.cap) * 16);
    q.n_out = reinterpret_cast<int*>(p);
    p += 256;
  }
  return det_launch(x, fr->data(), cfgs, n, x.stream, ev);
}

static double ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  return hipEventElapsedTime(&ms, a, b) == hipSuccess ? double(ms) : 0.0;
}

}  // namespace de

using namespace de;

extern "C" {

void de_detector_default_cfg(de_detector_cfg* c) {
  if (!c) return;
  // DetectorConfig (preproc.hpp:132-145)
  c->grad_window = 3;
  c->cell = 32;
  c->per_cell = 3;
  c->min_response = 1.0f;
  c->min_local_std = 2.0f;
  c->border = 8;
  c->subpixel = 1;
  c->fast_threshold = 8;
  c->nms_radius = 3;
}

size_t de_detect_max_keypoints(int W, int H, const de_detector_cfg* cfg) {
  if (!cfg || W <= 0 || H <= 0) return 0;
  return det_max_keypoints(*cfg, W, H);
}

int de_detect_keypoints_fast(de_ctx* c, const uint8_t* img, int W, int H,
                             const de_detector_cfg* cfg, de_keypoint* out, int cap, int* n_out,
                             de_detect_profile* prof) {
  if (!c || !cfg || !n_out) return fail(DE_ERR_ARG, "detect: null ctx, cfg or n_out");
  if (int rc = det_validate(*cfg, W, H)) return rc;
  *n_out = 0;
  if (prof) std::memset(prof, 0, sizeof(*prof));
  if (W == 0 || H == 0) return DE_OK;
  if (!img) return fail(DE_ERR_ARG, "detect: null image");
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  if (prof)
    for (auto& e : ev) DE_HIP(hipEventCreate(&e));
  std::vector<de_detect_frame> fr;
  int rc = det_host_run(x, &img, &W, &H, cfg, 1, &fr, prof ? ev : nullptr);
  int n = 0, counts[2] = {0, 0};
  if (!rc) rc = det_counts(x, fr.data(), cfg, 0, counts);
  if (!rc) {
    if (hipMemcpyAsync(&n, fr[0].n_out, sizeof(int), hipMemcpyDeviceToHost, x.stream) != hipSuccess ||
        hipStreamSynchronize(x.stream) != hipSuccess)
      rc = fail(DE_ERR_HIP, "detect: readback failed");
  }
  if (!rc && n > 0) {
    if (n > cap || !out) {
      rc = fail(DE_ERR_ARG, "detect: %d keypoints exceed the output capacity %d", n, cap);
    } else if (hipMemcpyAsync(out, fr[0].out, size_t(n) * sizeof(de_keypoint),
                              hipMemcpyDeviceToHost, x.stream) != hipSuccess ||
               hipStreamSynchronize(x.stream) != hipSuccess) {
      rc = fail(DE_ERR_HIP, "detect: readback failed");
    }
  }
  *n_out = n;
  if (prof && !rc) {
    prof->fast_ms = ev_ms(ev[0], ev[1]);
    prof->nms_ms = ev_ms(ev[1], ev[2]);
    prof->refine_ms = ev_ms(ev[2], ev[3]);
    prof->candidates = counts[0];
    prof->suppressed = counts[1];
  }
  if (prof)
    for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
}

int de_detect_keypoints_fast_batch_dev(de_ctx* c, const de_detect_frame* frames,
                                       const de_detector_cfg* cfgs, int n, void* stream) {
  if (!c || !frames || !cfgs || n < 1) return fail(DE_ERR_ARG, "detect batch: %d frames", n);
  for (int k = 0; k < n; ++k) {
    const de_detect_frame& q = frames[k];
    if (int rc = det_validate(cfgs[k], q.W, q.H)) return rc;
    if (q.W < 1 || q.H < 1 || !q.img || !q.out || !q.n_out)
      return fail(DE_ERR_ARG, "detect batch frame %d: empty image or null pointer", k);
    if (size_t(std::max(q.cap, 0)) < det_max_keypoints(cfgs[k], q.W, q.H))
      return fail(DE_ERR_ARG, "detect batch frame %d: cap %d below de_detect_max_keypoints %zu",
                  k, q.cap, det_max_keypoints(cfgs[k], q.W, q.H));
  }
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : x.stream;
  return det_launch(x, frames, cfgs, n, s, nullptr);
}

int de_fast_score_map(de_ctx* c, const uint8_t* img, int W, int H, int threshold,
                      int grad_window, uint8_t* corner, float* response) {
  if (!c || W < 0 || H < 0) return fail(DE_ERR_ARG, "fast map: bad arguments");
  if (W == 0 || H == 0) return DE_OK;
  if (!img || !corner || !response) return fail(DE_ERR_ARG, "fast map: null buffer");
  if (grad_window / 2 > 4096) return fail(DE_ERR_ARG, "fast map: grad_window %d", grad_window);
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  const size_t px = size_t(W) * size_t(H);
  if (int rc = x.det_io.ensure(a256(px) * 2 + a256(px * 4))) return rc;
  uint8_t* d_img = static_cast<uint8_t*>(x.det_io.ptr);
  uint8_t* d_c = d_img + a256(px);
  float* d_r = reinterpret_cast<float*>(d_c + a256(px));
  DE_HIP(hipMemcpyAsync(d_img, img, px, hipMemcpyHostToDevice, x.stream));
  hipLaunchKernelGGL(k_fast_map, dim3((W + 255) / 256, H), dim3(256), 0, x.stream, d_img, W, H,
                     threshold, std::max(1, grad_window / 2), d_c, d_r);
  DE_HIP(hipGetLastError());
  DE_HIP(hipMemcpyAsync(corner, d_c, px, hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipMemcpyAsync(response, d_r, px * 4, hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipStreamSynchronize(x.stream));
  return DE_OK;
}

int de_refine_disparity(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                        const de_keypoint* kl, int n_left, const de_keypoint* kr, int n_right,
                        de_match* matches, int n, int half) {
  (void)kr;   // right keypoints only bound the match's right index
  if (!c || W < 0 || H < 0 || n < 0 || n_left < 0 || n_right < 0 || half < 0)
    return fail(DE_ERR_ARG, "refine: bad arguments");
  if (n == 0 || W == 0 || H == 0) return DE_OK;
  if (!L || !R || !matches || (n_left > 0 && !kl))
    return fail(DE_ERR_ARG, "refine: null buffer");
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  const size_t px = size_t(W) * size_t(H);
  const size_t kb = size_t(std::max(n_left, 1)) * 16, mb = size_t(n) * sizeof(de_match);
  if (int rc = x.det_io.ensure(a256(px) * 2 + a256(kb) + a256(mb))) return rc;
  uint8_t* p = static_cast<uint8_t*>(x.det_io.ptr);
  uint8_t* dL = p;
  uint8_t* dR = p + a256(px);
  float4* dk = reinterpret_cast<float4*>(dR + a256(px));
  de_match* dm = reinterpret_cast<de_match*>(reinterpret_cast<uint8_t*>(dk) + a256(kb));
  DE_HIP(hipMemcpyAsync(dL, L, px, hipMemcpyHostToDevice, x.stream));
  DE_HIP(hipMemcpyAsync(dR, R, px, hipMemcpyHostToDevice, x.stream));
  if (n_left > 0) DE_HIP(hipMemcpyAsync(dk, kl, size_t(n_left) * 16, hipMemcpyHostToDevice, x.stream));
  DE_HIP(hipMemcpyAsync(dm, matches, mb, hipMemcpyHostToDevice, x.stream));
  hipLaunchKernelGGL(k_refine, dim3((n + 255) / 256), dim3(256), 0, x.stream, dL, dR, W, H, dk,
                     n_left, n_right, half, dm, n);
  DE_HIP(hipGetLastError());
  DE_HIP(hipMemcpyAsync(matches, dm, mb, hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipStreamSynchronize(x.stream));
  return DE_OK;
}

int de_census_at_keypoints(de_ctx* c, const uint8_t* img, int W, int H, int hw, int hh,
                           const de_keypoint* kp, int n, uint64_t* desc) {
  if (!c || W < 0 || H < 0 || n < 0 || hw < 0 || hh < 0)
    return fail(DE_ERR_ARG, "census at keypoints: bad arguments");
  if (n == 0) return DE_OK;
  if (!kp || !desc || (size_t(W) * size_t(H) > 0 && !img))
    return fail(DE_ERR_ARG, "census at keypoints: null buffer");
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  const size_t px = std::max<size_t>(size_t(W) * size_t(H), 1);
  if (int rc = x.det_io.ensure(a256(px) + a256(size_t(n) * 16) + a256(size_t(n) * 8))) return rc;
  uint8_t* d_img = static_cast<uint8_t*>(x.det_io.ptr);
  float4* d_kp = reinterpret_cast<float4*>(d_img + a256(px));
  uint64_t* d_desc = reinterpret_cast<uint64_t*>(reinterpret_cast<uint8_t*>(d_kp) + a256(size_t(n) * 16));
  if (size_t(W) * size_t(H) > 0)
    DE_HIP(hipMemcpyAsync(d_img, img, size_t(W) * H, hipMemcpyHostToDevice, x.stream));
  DE_HIP(hipMemcpyAsync(d_kp, kp, size_t(n) * 16, hipMemcpyHostToDevice, x.stream));
  hipLaunchKernelGGL(k_census_kp, dim3((n + 255) / 256), dim3(256), 0, x.stream, d_img, W, H, hw,
                     hh, d_kp, nullptr, n, d_desc);
  DE_HIP(hipGetLastError());
  DE_HIP(hipMemcpyAsync(desc, d_desc, size_t(n) * 8, hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipStreamSynchronize(x.stream));
  return DE_OK;
}

int de_sparse_pipeline(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                       const de_detector_cfg* det_left, const de_detector_cfg* det_right,
                       int census_hw, int census_hh, const de_match_cfg* mcfg, float min_margin,
                       int subpixel, de_keypoint* kl, int kl_cap, int* n_kl, de_keypoint* kr,
                       int kr_cap, int* n_kr, de_match* out, float* records, int out_cap,
                       int* n_out, de_match_stats* stats) {
  if (!c || !det_left || !mcfg || !n_kl || !n_kr || !n_out)
    return fail(DE_ERR_ARG, "pipeline: null argument");
  const de_detector_cfg cfgs[2] = {*det_left, det_right ? *det_right : *det_left};
  for (const auto& q : cfgs)
    if (int rc = det_validate(q, W, H)) return rc;
  if (census_hw < 0 || census_hh < 0) return fail(DE_ERR_ARG, "pipeline: census window");
  *n_kl = *n_kr = *n_out = 0;
  if (W == 0 || H == 0) return DE_OK;
  if (!L || !R) return fail(DE_ERR_ARG, "pipeline: null image");
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  const uint8_t* imgs[2] = {L, R};
  const int Ws[2] = {W, W}, Hs[2] = {H, H};
  std::vector<de_detect_frame> fr;
  if (int rc = det_host_run(x, imgs, Ws, Hs, cfgs, 2, &fr, nullptr)) return rc;
  int nk[2] = {0, 0};
  DE_HIP(hipMemcpyAsync(&nk[0], fr[0].n_out, sizeof(int), hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipMemcpyAsync(&nk[1], fr[1].n_out, sizeof(int), hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipStreamSynchronize(x.stream));
  const int nl = nk[0], nr = nk[1];
  if (nl > kl_cap || nr > kr_cap || (nl > 0 && !kl) || (nr > 0 && !kr))
    return fail(DE_ERR_ARG, "pipeline: %d / %d keypoints exceed the capacities %d / %d", nl, nr,
                kl_cap, kr_cap);
  if (int rc = sp_validate_cfg(*mcfg, nl, nr)) return rc;
  const int m = std::max(1, std::min(nl, nr));
  if (std::min(nl, nr) > 0 && (std::min(nl, nr) > out_cap || !out))
    return fail(DE_ERR_ARG, "pipeline: output capacity %d below %d", out_cap, std::min(nl, nr));
  // pipe buffer: descriptors, raw matches, finished matches, records, counts, stats
  const size_t pb = a256(size_t(std::max(nl, 1)) * 8) + a256(size_t(std::max(nr, 1)) * 8) +
                    a256(size_t(m) * sizeof(de_match)) * 2 + a256(size_t(m) * 16) + 512 +
                    a256(sizeof(de_match_stats));
  if (int rc = x.pipe_io.ensure(pb)) return rc;
  uint8_t* p = static_cast<uint8_t*>(x.pipe_io.ptr);
  uint64_t* dl = reinterpret_cast<uint64_t*>(p);
  p += a256(size_t(std::max(nl, 1)) * 8);
  uint64_t* dr = reinterpret_cast<uint64_t*>(p);
  p += a256(size_t(std::max(nr, 1)) * 8);
  de_match* raw = reinterpret_cast<de_match*>(p);
  p += a256(size_t(m) * sizeof(de_match));
  de_match* fin = reinterpret_cast<de_match*>(p);
  p += a256(size_t(m) * sizeof(de_match));
  float4* rec = reinterpret_cast<float4*>(p);
  p += a256(size_t(m) * 16);
  int* n_raw = reinterpret_cast<int*>(p);
  int* n_fin = reinterpret_cast<int*>(p + 256);
  p += 512;
  de_match_stats* d_st = reinterpret_cast<de_match_stats*>(p);
  if (nl > 0)
    hipLaunchKernelGGL(k_census_kp, dim3((nl + 255) / 256), dim3(256), 0, x.stream, fr[0].img, W,
                       H, census_hw, census_hh, reinterpret_cast<const float4*>(fr[0].out),
                       nullptr, nl, dl);
  if (nr > 0)
    hipLaunchKernelGGL(k_census_kp, dim3((nr + 255) / 256), dim3(256), 0, x.stream, fr[1].img, W,
                       H, census_hw, census_hh, reinterpret_cast<const float4*>(fr[1].out),
                       nullptr, nr, dr);
  DE_HIP(hipGetLastError());
  if (int rc = sparse_match_dev(x, fr[0].out, dl, nl, fr[1].out, dr, nr, *mcfg, raw, n_raw, d_st,
                                x.stream))
    return rc;
  hipLaunchKernelGGL(k_pipe_finish, dim3(1), dim3(kFinThreads), 0, x.stream, fr[0].img,
                     fr[1].img, W, H, reinterpret_cast<const float4*>(fr[0].out), nl, nr, raw,
                     n_raw, min_margin, subpixel ? 1 : 0, 3, fin, rec, n_fin);
  DE_HIP(hipGetLastError());
  int nf = 0;
  de_match_stats st;
  DE_HIP(hipMemcpyAsync(&nf, n_fin, sizeof(int), hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipMemcpyAsync(&st, d_st, sizeof(st), hipMemcpyDeviceToHost, x.stream));
  if (nl > 0) DE_HIP(hipMemcpyAsync(kl, fr[0].out, size_t(nl) * 16, hipMemcpyDeviceToHost, x.stream));
  if (nr > 0) DE_HIP(hipMemcpyAsync(kr, fr[1].out, size_t(nr) * 16, hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipStreamSynchronize(x.stream));
  if (nf > 0) {
    DE_HIP(hipMemcpyAsync(out, fin, size_t(nf) * sizeof(de_match), hipMemcpyDeviceToHost, x.stream));
    if (records)
      DE_HIP(hipMemcpyAsync(records, rec, size_t(nf) * 16, hipMemcpyDeviceToHost, x.stream));
    DE_HIP(hipStreamSynchronize(x.stream));
  }
  *n_kl = nl;
  *n_kr = nr;
  *n_out = nf;
  if (stats) *stats = st;
  return DE_OK;
}

}  // extern "C"
