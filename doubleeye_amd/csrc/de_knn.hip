// Keypoint-graph belief propagation on gfx950: labels along the epipolar row
// by census Hamming distance, a symmetric k-nearest-neighbour graph, min-sum
// loopy BP with a truncated-linear disparity-difference term, top-2 labels.
// The reference has no such solver (its sparse matcher is bipartite,
// core/src/match.cpp); the definition followed here, step for step and in the
// same floating-point order, is:
//   labels   keypoint i at (x, y): d in [0, dmax] with x - d >= hw;
//            h(d) = popcount(censusL(x, y) ^ censusR(x - d, y)); the n_cand
//            smallest by (h, d); unary cost c = w_data * h.
//   graph    j ~ i when j is among the k nearest of i or i among the k nearest
//            of j (squared distance, ties by index); lists ascending.
//   BP       B_i(l) = c_i(l) + sum_{k in N(i)} m_{k->i}(l), summed in ascending
//            k; m_{i->j}(q) = min_l [(B_i(l) - m_{j->i}(l)) + lambda *
//            min(|d_i(l) - d_j(q)|, trunc)], then minus its minimum over q; all
//            messages of an iteration from the previous ones; zero at start; a
//            node without labels sends zeros.
//   output   best and second label by (B, l).
//
// Launches (each over a batch of frames; items are keypoints unless noted):
//   census plane of each right image (the dense path's LDS-tiled kernel)
//   k_kb_cand   one wave per keypoint: left census by ballot, Hamming of the
//               row's right census words in lanes, n_cand rounds of wave min
//   k_kb_knn    one wave per keypoint: distances to all keypoints of the frame,
//               per-lane sorted lists, k rounds of wave min; appends i to the
//               reverse list of each neighbour
//   k_kb_adj    one wave per keypoint: union of both lists, deduplicated and
//               ordered by rank counting
//   k_kb_scan   one workgroup per frame: CSR offsets
//   k_kb_csr    one wave per keypoint: CSR lists, reverse-edge index by binary
//               search, messages zeroed
//   k_kb_iter   one wave per keypoint per iteration: beliefs in lanes 0..7,
//               outgoing messages in lanes (neighbour, target label)
//   k_kb_out    one wave per keypoint: final beliefs and the top-2
// Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "de_api_internal.h"
#include "de_common.h"

namespace de {

namespace {

constexpr int kLab = 8;          // labels per keypoint (max n_cand)
constexpr int kMaxK = 8;         // max k_nn
constexpr int kMaxRev = 56;      // reverse-list capacity per keypoint
constexpr int kMaxDeg = 64;      // neighbours per keypoint (one lane each)
constexpr int kWavesPerBlock = 4;
constexpr int kScanThreads = 1024;

enum KbFlag { kFlagRevOverflow = 1, kFlagAsymmetric = 2, kFlagKeypoint = 4 };

struct KbDesc {
  const uint8_t* L;
  const uint8_t* R;
  const int2* kp;          // keypoints (x, y)
  const uint64_t* crs;     // right census plane, W*H (null when labels are given)
  int W, H, n;
  int hw, hh, dmax, n_cand, k_nn, iters, trunc;
  float w_data, lambda;
  // graph (built here, or given)
  int32_t* cand_d;         // [n][8]
  float* cand_c;           // [n][8]
  int32_t* ncand;          // [n]
  int32_t* knn;            // [n][8]
  int32_t* rdeg;           // [n]
  int32_t* rlist;          // [n][kMaxRev]
  int32_t* deg;            // [n]
  int32_t* adj_tmp;        // [n][kMaxDeg]
  int32_t* off;            // [n + 1]
  int32_t* adj;            // [n * kMaxDeg]
  int32_t* rev;            // [n * kMaxDeg]
  float* msg[2];           // [n * kMaxDeg][8]
  int32_t* flag;           // error bits
  // outputs
  de_knn_out* out;
  float* beliefs;          // nullable [n][8]
};

struct KbBatch {
  const KbDesc* frames;
  const int* start;        // [n + 1] prefix of keypoints
  int n;
};

__device__ __forceinline__ int frame_of(const KbBatch& b, int item) {
  int lo = 0, hi = b.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b.start[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// The wave's item (keypoint across the batch) and lane.
struct WaveItem {
  int item, lane;
};
__device__ __forceinline__ WaveItem wave_item() {
  WaveItem w;
  w.item = __builtin_amdgcn_readfirstlane(int(blockIdx.x) * kWavesPerBlock +
                                          int(threadIdx.x / kWave));
  w.lane = int(threadIdx.x & (kWave - 1));
  return w;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t w = __shfl_xor(v, o);
    v = w < v ? w : v;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
  for (int o = 1; o < kWave; o <<= 1) {
    const uint64_t w = __shfl_xor(v, o);
    v = w < v ? w : v;
  }
  return v;
}

// Beliefs of node i in lanes 0..7 (label = lane), summing incoming messages in
// ascending neighbour order. rv: the reverse-edge index held by lane s.
__device__ float node_belief(const KbDesc& f, const float* __restrict__ m, int i, int o0,
                             int dg, int rv, int lane, bool lab_ok) {
  float v = lab_ok ? f.cand_c[i * kLab + (lane & 7)] : 0.f;
  for (int s0 = 0; s0 < dg; s0 += 8) {
    float t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = __shfl(rv, s0 + u);
      t[u] = (s0 + u < dg && lab_ok) ? m[size_t(e) * kLab + (lane & 7)] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (s0 + u < dg) v = v + t[u];
  }
  return v;
}

// ---------------------------------------------------------------------------

// One thread per frame: clear the error bits before any stage can set them.
__global__ void k_kb_init(const KbDesc* __restrict__ frames, int n) {
  const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (t < n) *frames[t].flag = 0;
}

__global__ __launch_bounds__(kWavesPerBlock * kWave) void k_kb_cand(KbBatch b, int total) {
  const WaveItem w = wave_item();
  if (w.item >= total) return;
  const int fi = frame_of(b, w.item);
  const KbDesc& f = b.frames[fi];
  const int i = w.item - b.start[fi];
  const int lane = w.lane;
  if (lane == 0) f.rdeg[i] = 0;
  const int2 p = f.kp[i];
  const int W = f.W;
  if (p.x < f.hw || p.x >= W - f.hw || p.y < f.hh || p.y >= f.H - f.hh) {
    // census window outside the image: no labels, reported through the flag
    if (lane == 0) {
      f.ncand[i] = 0;
      atomicOr(f.flag, int(kFlagKeypoint));
    }
    if (lane < kLab) {
      f.cand_d[i * kLab + lane] = -1;
      f.cand_c[i * kLab + lane] = 0.f;
    }
    return;
  }
  // left census at (x, y): lane b computes bit b (dy-major, dx-minor, centre skipped)
  const int ww = 2 * f.hw + 1, nb = ww * (2 * f.hh + 1) - 1;
  const int centre = f.hh * ww + f.hw;
  const uint8_t c = f.L[size_t(p.y) * W + p.x];
  bool bit = false;
  if (lane < nb) {
    const int idx = lane < centre ? lane : lane + 1;
    const int dy = idx / ww - f.hh, dx = idx % ww - f.hw;
    bit = f.L[size_t(p.y + dy) * W + (p.x + dx)] < c;
  }
  const uint64_t cl = __ballot(bit);
  // Hamming distance for d = lane + 64 r
  uint32_t key[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int d = lane + kWave * r;
    key[r] = 0xFFFFFFFFu;
    if (d <= f.dmax && p.x - d >= f.hw) {
      const uint64_t cr = f.crs[size_t(p.y) * W + (p.x - d)];
      key[r] = (uint32_t(__popcll(cl ^ cr)) << 8) | uint32_t(d);
    }
  }
  uint32_t thr = 0;
  int m = 0;
  for (; m < f.n_cand; ++m) {
    uint32_t mine = 0xFFFFFFFFu;
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if ((m == 0 || key[r] > thr) && key[r] < mine) mine = key[r];
    const uint32_t best = wave_min_u32(mine);
    if (best == 0xFFFFFFFFu) break;
    if (lane == 0) {
      f.cand_d[i * kLab + m] = int(best & 255u);
      f.cand_c[i * kLab + m] = f.w_data * float(best >> 8);
    }
    thr = best;
  }
  if (lane == 0) f.ncand[i] = m;
  if (lane >= m && lane < kLab) {
    f.cand_d[i * kLab + lane] = -1;
    f.cand_c[i * kLab + lane] = 0.f;
  }
}

__global__ __launch_bounds__(kWavesPerBlock * kWave) void k_kb_knn(KbBatch b, int total) {
  const WaveItem w = wave_item();
  if (w.item >= total) return;
  const int fi = frame_of(b, w.item);
  const KbDesc& f = b.frames[fi];
  const int i = w.item - b.start[fi];
  const int lane = w.lane, n = f.n;
  const int kk = min(f.k_nn, n - 1);
  if (kk <= 0) return;
  const int2 p = f.kp[i];
  // per-lane ascending list of the kMaxK best (dist2 << 32 | j)
  uint64_t lst[kMaxK];
#pragma unroll
  for (int t = 0; t < kMaxK; ++t) lst[t] = ~0ull;
  for (int j = lane; j < n; j += kWave) {
    if (j == i) continue;
    const int2 q = f.kp[j];
    const int64_t dx = q.x - p.x, dy = q.y - p.y;
    const uint64_t key = (uint64_t(dx * dx + dy * dy) << 32) | uint32_t(j);
    if (key < lst[kMaxK - 1]) {
      uint64_t v = key;
#pragma unroll
      for (int t = 0; t < kMaxK; ++t) {   // insertion: keep ascending
        const uint64_t lo = v < lst[t] ? v : lst[t];
        const uint64_t hi = v < lst[t] ? lst[t] : v;
        lst[t] = lo;
        v = hi;
      }
    }
  }
  for (int t = 0; t < kk; ++t) {
    const uint64_t best = wave_min_u64(lst[0]);
    if (lst[0] == best) {   // the owning lane pops its head
#pragma unroll
      for (int u = 0; u < kMaxK - 1; ++u) lst[u] = lst[u + 1];
      lst[kMaxK - 1] = ~0ull;
    }
    const int j = int(uint32_t(best));
    if (lane == 0) {
      f.knn[i * kMaxK + t] = j;
      const int slot = atomicAdd(&f.rdeg[j], 1);
      if (slot < kMaxRev) f.rlist[j * kMaxRev + slot] = i;
      else atomicOr(f.flag, int(kFlagRevOverflow));
    }
  }
}

__global__ __launch_bounds__(kWavesPerBlock * kWave) void k_kb_adj(KbBatch b, int total) {
  const WaveItem w = wave_item();
  if (w.item >= total) return;
  const int fi = frame_of(b, w.item);
  const KbDesc& f = b.frames[fi];
  const int i = w.item - b.start[fi];
  const int lane = w.lane;
  const int kk = max(0, min(f.k_nn, f.n - 1));
  const int nr = min(f.rdeg[i], kMaxRev);
  int v = 0x7FFFFFFF;
  if (lane < kk) v = f.knn[i * kMaxK + lane];
  else if (lane < kk + nr) v = f.rlist[i * kMaxRev + (lane - kk)];
  const bool valid = lane < kk + nr;
  // keep the first occurrence of each value; its position is the number of
  // distinct values below it
  const int cnt = kk + nr;
  bool first = valid;
  for (int t = 0; t < cnt; ++t) {
    const int u = __shfl(v, t);
    if (t < lane && u == v) first = false;
  }
  int pos = 0;
  for (int t = 0; t < cnt; ++t) {
    const int u = __shfl(v, t);
    const int ft = __shfl(int(first), t);
    if (ft && u < v) pos += 1;
  }
  const int dg = __popcll(__ballot(first));
  if (first) f.adj_tmp[i * kMaxDeg + pos] = v;
  if (lane == 0) f.deg[i] = dg;
}

// One workgroup per frame: off[] = exclusive prefix of deg[].
__global__ __launch_bounds__(kScanThreads) void k_kb_scan(const KbDesc* __restrict__ frames) {
  __shared__ int sh[kScanThreads / kWave];
  const KbDesc& f = frames[blockIdx.x];
  const int n = f.n, tid = threadIdx.x;
  const int ch = (n + kScanThreads - 1) / kScanThreads;
  const int r0 = min(n, tid * ch), r1 = min(n, r0 + ch);
  int run = 0;
  for (int r = r0; r < r1; ++r) run += f.deg[r];
  const int lane = tid & (kWave - 1), wv = tid / kWave;
  int x = run;
  for (int o = 1; o < kWave; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) sh[wv] = x;
  __syncthreads();
  int before = 0, all = 0;
  for (int k = 0; k < kScanThreads / kWave; ++k) {
    if (k < wv) before += sh[k];
    all += sh[k];
  }
  int acc = before + x - run;
  for (int r = r0; r < r1; ++r) {
    f.off[r] = acc;
    acc += f.deg[r];
  }
  if (tid == 0) f.off[n] = all;
}

// CSR lists (from adj_tmp when `copy`), reverse-edge index, zero messages.
__global__ __launch_bounds__(kWavesPerBlock * kWave) void k_kb_csr(KbBatch b, int total,
                                                                   int copy) {
  const WaveItem w = wave_item();
  if (w.item >= total) return;
  const int fi = frame_of(b, w.item);
  const KbDesc& f = b.frames[fi];
  const int i = w.item - b.start[fi];
  const int lane = w.lane;
  const int o0 = f.off[i], dg = f.off[i + 1] - o0;
  if (lane < dg) {
    int j;
    if (copy) {
      j = f.adj_tmp[i * kMaxDeg + lane];
      f.adj[o0 + lane] = j;
    } else {
      j = f.adj[o0 + lane];
    }
    // position of i in j's ascending list
    const int* lj = copy ? f.adj_tmp + j * kMaxDeg : f.adj + f.off[j];
    int lo = 0, hi = copy ? f.deg[j] : f.off[j + 1] - f.off[j];
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (lj[mid] < i) lo = mid + 1; else hi = mid;
    }
    const int hij = copy ? f.deg[j] : f.off[j + 1] - f.off[j];
    if (lo < hij && lj[lo] == i) f.rev[o0 + lane] = f.off[j] + lo;
    else {
      f.rev[o0 + lane] = o0 + lane;   // keeps later reads in bounds
      atomicOr(f.flag, int(kFlagAsymmetric));
    }
  }
  for (int t = lane; t < dg * kLab; t += kWave) f.msg[0][size_t(o0) * kLab + t] = 0.f;
}

__global__ __launch_bounds__(kWavesPerBlock * kWave) void k_kb_iter(KbBatch b, int total,
                                                                    int src) {
  const WaveItem w = wave_item();
  if (w.item >= total) return;
  const int fi = frame_of(b, w.item);
  const KbDesc& f = b.frames[fi];
  const int i = w.item - b.start[fi];
  const int lane = w.lane;
  const float* __restrict__ mo = f.msg[src];
  float* __restrict__ mn = f.msg[src ^ 1];
  const int o0 = f.off[i], dg = f.off[i + 1] - o0;
  const int ni = f.ncand[i];
  const int rv = lane < dg ? f.rev[o0 + lane] : 0;
  const float B = node_belief(f, mo, i, o0, dg, rv, lane, lane < ni);
  const int di = lane < kLab ? f.cand_d[i * kLab + lane] : 0;
  float Bl[kLab];
  int dl[kLab];
#pragma unroll
  for (int l = 0; l < kLab; ++l) {   // every lane active: plain broadcasts
    Bl[l] = __shfl(B, l);
    dl[l] = __shfl(di, l);
  }
  // lanes (g, q): neighbour s = s0 + g, target label q
  const int g = lane >> 3, q = lane & 7;
  for (int s0 = 0; s0 < dg; s0 += 8) {
    const int s = s0 + g;
    const bool have = s < dg;
    const int e = o0 + (have ? s : 0);
    const int j = have ? f.adj[e] : 0;
    const int nj = have ? f.ncand[j] : 0;
    const int back_e = __shfl(rv, have ? s : 0);
    float best = INFINITY;
    const bool live = have && q < nj && ni > 0;
    if (live) {
      const int dj = f.cand_d[j * kLab + q];
      float bk[kLab];
#pragma unroll
      for (int l = 0; l < kLab; ++l) bk[l] = l < ni ? mo[size_t(back_e) * kLab + l] : 0.f;
#pragma unroll
      for (int l = 0; l < kLab; ++l) {
        if (l < ni) {
          const int dd = abs(dl[l] - dj);
          const float v = (Bl[l] - bk[l]) + f.lambda * float(min(dd, f.trunc));
          best = v < best ? v : best;
        }
      }
    }
    // minimum over the group's valid targets
    float mnv = live ? best : INFINITY;
    for (int o = 1; o < 8; o <<= 1) {
      const float t = __shfl_xor(mnv, o, 8);
      mnv = t < mnv ? t : mnv;
    }
    if (have) mn[size_t(e) * kLab + q] = live ? best - mnv : 0.f;
  }
}

__global__ __launch_bounds__(kWavesPerBlock * kWave) void k_kb_out(KbBatch b, int total,
                                                                   int src) {
  const WaveItem w = wave_item();
  if (w.item >= total) return;
  const int fi = frame_of(b, w.item);
  const KbDesc& f = b.frames[fi];
  const int i = w.item - b.start[fi];
  const int lane = w.lane;
  const int o0 = f.off[i], dg = f.off[i + 1] - o0;
  const int ni = f.ncand[i];
  const int rv = lane < dg ? f.rev[o0 + lane] : 0;
  const float B = node_belief(f, f.msg[src], i, o0, dg, rv, lane, lane < ni);
  if (f.beliefs && lane < kLab) f.beliefs[i * kLab + lane] = lane < ni ? B : INFINITY;
  float Bl[kLab];
#pragma unroll
  for (int l = 0; l < kLab; ++l) Bl[l] = __shfl(B, l);   // every lane active
  if (lane == 0) {
    int l1 = -1, l2 = -1;
    float b1 = INFINITY, b2 = INFINITY;
    for (int l = 0; l < ni; ++l) {
      const float v = Bl[l];
      if (l1 < 0 || v < b1) { l2 = l1; b2 = b1; l1 = l; b1 = v; }
      else if (l2 < 0 || v < b2) { l2 = l; b2 = v; }
    }
    de_knn_out o;
    o.d1 = l1 >= 0 ? f.cand_d[i * kLab + l1] : -1;
    o.d2 = l2 >= 0 ? f.cand_d[i * kLab + l2] : -1;
    o.c1 = b1;
    o.c2 = b2;
    f.out[i] = o;
  }
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

size_t kb_scratch_bytes(int W, int H, int n, bool census) {
  const size_t N = size_t(std::max(n, 1));
  size_t b = 0;
  if (census) b += align256(size_t(W) * H * 8);
  b += align256(N * kLab * 4) * 2 + align256(N * 4) + align256(N * kMaxK * 4) +
       align256(N * 4) + align256(N * kMaxRev * 4) + align256(N * 4) +
       align256(N * kMaxDeg * 4) + align256((N + 1) * 4) + align256(N * kMaxDeg * 4) * 2 +
       align256(N * kMaxDeg * kLab * 4) * 2 + align256(4);
  return b;
}

void kb_carve(void* base, int W, int H, int n, bool census, KbDesc* d) {
  uint8_t* p = static_cast<uint8_t*>(base);
  auto take = [&](size_t bytes) { void* r = p; p += align256(bytes); return r; };
  const size_t N = size_t(std::max(n, 1));
  d->crs = census ? static_cast<const uint64_t*>(take(size_t(W) * H * 8)) : nullptr;
  d->cand_d = static_cast<int32_t*>(take(N * kLab * 4));
  d->cand_c = static_cast<float*>(take(N * kLab * 4));
  d->ncand = static_cast<int32_t*>(take(N * 4));
  d->knn = static_cast<int32_t*>(take(N * kMaxK * 4));
  d->rdeg = static_cast<int32_t*>(take(N * 4));
  d->rlist = static_cast<int32_t*>(take(N * kMaxRev * 4));
  d->deg = static_cast<int32_t*>(take(N * 4));
  d->adj_tmp = static_cast<int32_t*>(take(N * kMaxDeg * 4));
  d->off = static_cast<int32_t*>(take((N + 1) * 4));
  d->adj = static_cast<int32_t*>(take(N * kMaxDeg * 4));
  d->rev = static_cast<int32_t*>(take(N * kMaxDeg * 4));
  d->msg[0] = static_cast<float*>(take(N * kMaxDeg * kLab * 4));
  d->msg[1] = static_cast<float*>(take(N * kMaxDeg * kLab * 4));
  d->flag = static_cast<int32_t*>(take(4));
}

int kb_validate(const de_knn_cfg& c, int W, int H, int n) {
  if (n < 0) return fail(DE_ERR_ARG, "knn: %d keypoints", n);
  if (W < 1 || H < 1) return fail(DE_ERR_ARG, "knn: image %dx%d", W, H);
  if (!((c.census_hw == 3 || c.census_hw == 4) && c.census_hh == 3))
    return fail(DE_ERR_ARG, "knn: census window %dx%d (7x7 or 9x7)", 2 * c.census_hw + 1,
                2 * c.census_hh + 1);
  if (c.dmax < 0 || c.dmax > 255) return fail(DE_ERR_ARG, "knn: dmax %d not in [0, 255]", c.dmax);
  if (c.n_cand < 1 || c.n_cand > kLab)
    return fail(DE_ERR_ARG, "knn: n_cand %d not in [1, %d]", c.n_cand, kLab);
  if (c.k_nn < 0 || c.k_nn > kMaxK) return fail(DE_ERR_ARG, "knn: k_nn %d not in [0, %d]", c.k_nn, kMaxK);
  if (c.iters < 0) return fail(DE_ERR_ARG, "knn: iters %d < 0", c.iters);
  if (c.trunc < 0) return fail(DE_ERR_ARG, "knn: trunc %d < 0", c.trunc);
  return DE_OK;
}

void kb_fill_cfg(const de_knn_cfg& c, KbDesc* d) {
  d->hw = c.census_hw;
  d->hh = c.census_hh;
  d->dmax = c.dmax;
  d->n_cand = c.n_cand;
  d->k_nn = c.k_nn;
  d->iters = c.iters;
  d->trunc = c.trunc;
  d->w_data = c.w_data;
  d->lambda = c.lambda;
}

int blocks_for(int total) { return (total + kWavesPerBlock - 1) / kWavesPerBlock; }

}  // namespace

// Stages over a batch whose descriptors are at d_frames: build = labels and
// graph from the images (else labels and CSR lists are given in scratch).
static int kb_launch(const std::vector<KbDesc>& fr, const KbDesc* d_frames,
                     const int* d_start, int total, bool build, int iters, hipStream_t s) {
  const dim3 wb(kWavesPerBlock * kWave);
  KbBatch b;
  b.frames = d_frames;
  b.start = d_start;
  b.n = int(fr.size());
  hipLaunchKernelGGL(k_kb_init, dim3((b.n + 255) / 256), dim3(256), 0, s, d_frames, b.n);
  DE_HIP(hipGetLastError());
  if (build) {
    for (const KbDesc& f : fr)
      if (f.n > 0)
        DE_HIP(launch_census_plane(f.R, f.W, f.H, f.hw, f.hh, const_cast<uint64_t*>(f.crs), s));
  }
  if (total > 0 && build) {
    hipLaunchKernelGGL(k_kb_cand, dim3(blocks_for(total)), wb, 0, s, b, total);
    hipLaunchKernelGGL(k_kb_knn, dim3(blocks_for(total)), wb, 0, s, b, total);
    hipLaunchKernelGGL(k_kb_adj, dim3(blocks_for(total)), wb, 0, s, b, total);
    DE_HIP(hipGetLastError());
  }
  if (build) {
    hipLaunchKernelGGL(k_kb_scan, dim3(b.n), dim3(kScanThreads), 0, s, d_frames);
    DE_HIP(hipGetLastError());
  }
  if (total > 0) {
    hipLaunchKernelGGL(k_kb_csr, dim3(blocks_for(total)), wb, 0, s, b, total, build ? 1 : 0);
    for (int it = 0; it < iters; ++it)
      hipLaunchKernelGGL(k_kb_iter, dim3(blocks_for(total)), wb, 0, s, b, total, it & 1);
    hipLaunchKernelGGL(k_kb_out, dim3(blocks_for(total)), wb, 0, s, b, total, iters & 1);
    DE_HIP(hipGetLastError());
  }
  return DE_OK;
}


// Upload descriptors through the context's descriptor ring and launch.
static int kb_run(Ctx& x, const std::vector<KbDesc>& fr, bool build, int iters, hipStream_t s) {
  const int n = int(fr.size());
  std::vector<int> start(size_t(n) + 1, 0);
  for (int k = 0; k < n; ++k) start[size_t(k) + 1] = start[size_t(k)] + fr[size_t(k)].n;
  const size_t db = fr.size() * sizeof(KbDesc), sb = start.size() * sizeof(int);
  Ctx::Slot* slot;
  if (int rc = x.acquire_slot(db + sb, &slot)) return rc;
  std::memcpy(slot->host, fr.data(), db);
  std::memcpy(static_cast<uint8_t*>(slot->host) + db, start.data(), sb);
  DE_HIP(hipMemcpyAsync(slot->dev, slot->host, db + sb, hipMemcpyHostToDevice, s));
  const KbDesc* d_frames = static_cast<const KbDesc*>(slot->dev);
  const int* d_start = reinterpret_cast<const int*>(static_cast<uint8_t*>(slot->dev) + db);
  if (int rc = kb_launch(fr, d_frames, d_start, start[size_t(n)], build, iters, s)) return rc;
  DE_HIP(hipEventRecord(slot->done, s));
  slot->used = true;
  return DE_OK;
}

static int kb_flags_to_status(int flags) {
  if (flags & kFlagKeypoint)
    return fail(DE_ERR_ARG, "knn: a keypoint's census window leaves the image");
  if (flags & kFlagRevOverflow)
    return fail(DE_ERR_UNSUPPORTED, "knn: a keypoint is a nearest neighbour of more than %d "
                "others (duplicate positions?)", kMaxRev);
  if (flags & kFlagAsymmetric) return fail(DE_ERR_ARG, "knn: the graph is not symmetric");
  return DE_OK;
}

// Host-pointer forms: stage inputs in the context's io buffer, run one frame.
// mode 0: whole workload; 1: labels only; 2: graph only; 3: BP on a given graph.
static int kb_host(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                   const int32_t* kp, int n, const de_knn_cfg* cfg, int mode,
                   const int32_t* g_cand_d, const float* g_cand_c, const int32_t* g_ncand,
                   const int32_t* g_off, const int32_t* g_adj, de_knn_out* out, float* beliefs,
                   int32_t* o_cand_d, float* o_cand_c, int32_t* o_ncand, int32_t* o_off,
                   int32_t* o_adj, int adj_cap, int* n_edges) {
  if (!c || !cfg) return fail(DE_ERR_ARG, "knn: null ctx or cfg");
  const bool build = mode != 3;
  if (int rc = kb_validate(*cfg, build ? W : 1, build ? H : 1, n)) return rc;
  if (build && n > 0 && (!L || !R || !kp)) return fail(DE_ERR_ARG, "knn: null input");
  if (!build) {
    if (n > 0 && (!g_cand_d || !g_cand_c || !g_ncand || !g_off || (g_off[n] > 0 && !g_adj)))
      return fail(DE_ERR_ARG, "knn: null graph input");
    for (int i = 0; i < n; ++i) {
      const int dg = g_off[i + 1] - g_off[i];
      if (dg < 0 || dg > kMaxDeg)
        return fail(DE_ERR_ARG, "knn: node %d has %d neighbours (at most %d)", i, dg, kMaxDeg);
      if (g_ncand[i] < 0 || g_ncand[i] > kLab)
        return fail(DE_ERR_ARG, "knn: node %d has %d labels (at most %d)", i, g_ncand[i], kLab);
      for (int e = g_off[i]; e < g_off[i + 1]; ++e) {
        if (g_adj[e] < 0 || g_adj[e] >= n || g_adj[e] == i ||
            (e > g_off[i] && g_adj[e] <= g_adj[e - 1]))
          return fail(DE_ERR_ARG, "knn: node %d's neighbour list must be ascending, in range "
                      "and without itself", i);
      }
    }
    if (n > 0 && g_off[0] != 0) return fail(DE_ERR_ARG, "knn: off[0] must be 0");
  }
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  const size_t WH = build ? size_t(W) * H : 0;
  const size_t N = size_t(std::max(n, 1));
  const size_t io = align256(WH) * 2 + align256(N * 8) + align256(N * sizeof(de_knn_out)) +
                    align256(N * kLab * 4);
  if (int rc = x.kb_io.ensure(io)) return rc;
  if (int rc = x.kb_scratch.ensure(kb_scratch_bytes(W, H, n, build))) return rc;
  KbDesc d;
  std::memset(&d, 0, sizeof(d));
  kb_fill_cfg(*cfg, &d);
  d.W = build ? W : 1;
  d.H = build ? H : 1;
  d.n = n;
  kb_carve(x.kb_scratch.ptr, W, H, n, build, &d);
  uint8_t* p = static_cast<uint8_t*>(x.kb_io.ptr);
  if (build) {
    d.L = p;
    p += align256(WH);
    d.R = p;
    p += align256(WH);
    d.kp = reinterpret_cast<const int2*>(p);
    if (n > 0) {
      DE_HIP(hipMemcpyAsync(const_cast<uint8_t*>(d.L), L, WH, hipMemcpyHostToDevice, x.stream));
      DE_HIP(hipMemcpyAsync(const_cast<uint8_t*>(d.R), R, WH, hipMemcpyHostToDevice, x.stream));
      DE_HIP(hipMemcpyAsync(p, kp, size_t(n) * 8, hipMemcpyHostToDevice, x.stream));
    }
  } else if (n > 0) {
    DE_HIP(hipMemcpyAsync(d.cand_d, g_cand_d, size_t(n) * kLab * 4, hipMemcpyHostToDevice, x.stream));
    DE_HIP(hipMemcpyAsync(d.cand_c, g_cand_c, size_t(n) * kLab * 4, hipMemcpyHostToDevice, x.stream));
    DE_HIP(hipMemcpyAsync(d.ncand, g_ncand, size_t(n) * 4, hipMemcpyHostToDevice, x.stream));
    DE_HIP(hipMemcpyAsync(d.off, g_off, (size_t(n) + 1) * 4, hipMemcpyHostToDevice, x.stream));
    if (g_off[n] > 0)
      DE_HIP(hipMemcpyAsync(d.adj, g_adj, size_t(g_off[n]) * 4, hipMemcpyHostToDevice, x.stream));
  }
  p += align256(N * 8);
  d.out = reinterpret_cast<de_knn_out*>(p);
  p += align256(N * sizeof(de_knn_out));
  d.beliefs = beliefs ? reinterpret_cast<float*>(p) : nullptr;
  const int iters = (mode == 0 || mode == 3) ? cfg->iters : 0;
  std::vector<KbDesc> fr(1, d);
  if (n > 0 || mode == 2) {
    if (int rc = kb_run(x, fr, build, iters, x.stream)) return rc;
  }
  int flags = 0;
  if (n > 0) DE_HIP(hipMemcpyAsync(&flags, d.flag, 4, hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipStreamSynchronize(x.stream));
  if (int rc = kb_flags_to_status(flags)) return rc;
  if (n == 0) {
    if (o_off) o_off[0] = 0;
    if (n_edges) *n_edges = 0;
    return DE_OK;
  }
  if (out) DE_HIP(hipMemcpy(out, d.out, size_t(n) * sizeof(de_knn_out), hipMemcpyDeviceToHost));
  if (beliefs) DE_HIP(hipMemcpy(beliefs, d.beliefs, size_t(n) * kLab * 4, hipMemcpyDeviceToHost));
  if (o_cand_d) DE_HIP(hipMemcpy(o_cand_d, d.cand_d, size_t(n) * kLab * 4, hipMemcpyDeviceToHost));
  if (o_cand_c) DE_HIP(hipMemcpy(o_cand_c, d.cand_c, size_t(n) * kLab * 4, hipMemcpyDeviceToHost));
  if (o_ncand) DE_HIP(hipMemcpy(o_ncand, d.ncand, size_t(n) * 4, hipMemcpyDeviceToHost));
  if (o_off || o_adj || n_edges) {
    std::vector<int32_t> off(size_t(n) + 1);
    DE_HIP(hipMemcpy(off.data(), d.off, (size_t(n) + 1) * 4, hipMemcpyDeviceToHost));
    if (n_edges) *n_edges = off[size_t(n)];
    if (o_off) std::memcpy(o_off, off.data(), (size_t(n) + 1) * 4);
    if (o_adj) {
      if (off[size_t(n)] > adj_cap)
        return fail(DE_ERR_ARG, "knn: %d edges exceed adj capacity %d", off[size_t(n)], adj_cap);
      if (off[size_t(n)] > 0)
        DE_HIP(hipMemcpy(o_adj, d.adj, size_t(off[size_t(n)]) * 4, hipMemcpyDeviceToHost));
    }
  }
  return DE_OK;
}

size_t knn_scratch_bytes(int W, int H, int n) { return kb_scratch_bytes(W, H, n, true); }

int knn_validate(const de_knn_cfg& c, int W, int H, int n) { return kb_validate(c, W, H, n); }

int knn_launch_frames(Ctx& x, const de_knn_frame* frames, const de_knn_cfg* cfgs,
                      void* const* scratch, int n, hipStream_t s) {
  for (int k = 0; k < n; ++k) {
    const de_knn_frame& q = frames[k];
    if (int rc = kb_validate(cfgs[k], q.W, q.H, q.n)) return rc;
    if (cfgs[k].iters != cfgs[0].iters)
      return fail(DE_ERR_ARG, "knn batch frame %d: iters must match the first frame's", k);
    if (q.n > 0 && (!q.L || !q.R || !q.kp || !q.out))
      return fail(DE_ERR_ARG, "knn batch frame %d: null pointer", k);
  }
  std::vector<KbDesc> fr(static_cast<size_t>(n));
  for (int k = 0; k < n; ++k) {
    const de_knn_frame& q = frames[k];
    KbDesc& d = fr[size_t(k)];
    std::memset(&d, 0, sizeof(d));
    kb_fill_cfg(cfgs[k], &d);
    d.L = q.L;
    d.R = q.R;
    d.kp = reinterpret_cast<const int2*>(q.kp);
    d.W = q.W;
    d.H = q.H;
    d.n = q.n;
    d.out = q.out;
    kb_carve(scratch[k], q.W, q.H, q.n, true, &d);
  }
  return kb_run(x, fr, true, cfgs[0].iters, s);
}

}  // namespace de

using namespace de;

extern "C" {

void de_knn_default_cfg(de_knn_cfg* c) {
  if (!c) return;
  c->census_hw = 4;
  c->census_hh = 3;
  c->dmax = 127;
  c->n_cand = 8;
  c->k_nn = 8;
  c->iters = 10;
  c->w_data = 1.0f;
  c->lambda = 1.0f;
  c->trunc = 6;
}

int de_knn_bp(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H, const int32_t* kp,
              int n, const de_knn_cfg* cfg, de_knn_out* out) {
  if (!out && n > 0) return fail(DE_ERR_ARG, "de_knn_bp: null output");
  return kb_host(c, L, R, W, H, kp, n, cfg, 0, nullptr, nullptr, nullptr, nullptr, nullptr,
                 out, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr);
}

int de_knn_candidates(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                      const int32_t* kp, int n, const de_knn_cfg* cfg, int32_t* cand_d,
                      float* cand_c, int32_t* ncand) {
  return kb_host(c, L, R, W, H, kp, n, cfg, 1, nullptr, nullptr, nullptr, nullptr, nullptr,
                 nullptr, nullptr, cand_d, cand_c, ncand, nullptr, nullptr, 0, nullptr);
}

int de_knn_graph(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                 const int32_t* kp, int n, const de_knn_cfg* cfg, int32_t* off, int32_t* adj,
                 int adj_cap, int* n_edges) {
  return kb_host(c, L, R, W, H, kp, n, cfg, 2, nullptr, nullptr, nullptr, nullptr, nullptr,
                 nullptr, nullptr, nullptr, nullptr, nullptr, off, adj, adj_cap, n_edges);
}

int de_knn_bp_graph(de_ctx* c, int n, const int32_t* cand_d, const float* cand_c,
                    const int32_t* ncand, const int32_t* off, const int32_t* adj,
                    const de_knn_cfg* cfg, de_knn_out* out, float* beliefs) {
  if (!out && n > 0) return fail(DE_ERR_ARG, "de_knn_bp_graph: null output");
  return kb_host(c, nullptr, nullptr, 1, 1, nullptr, n, cfg, 3, cand_d, cand_c, ncand, off, adj,
                 out, beliefs, nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr);
}

int de_knn_bp_batch_dev(de_ctx* c, const de_knn_frame* frames, const de_knn_cfg* cfgs, int n,
                        void* stream) {
  if (!c || !frames || !cfgs || n < 1) return fail(DE_ERR_ARG, "knn batch: %d frames", n);
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : x.stream;
  std::vector<size_t> offs(static_cast<size_t>(n));
  size_t total = 0;
  for (int k = 0; k < n; ++k) {
    offs[size_t(k)] = total;
    total += align256(knn_scratch_bytes(frames[k].W, frames[k].H, frames[k].n));
  }
  if (int rc = x.kb_scratch.ensure(total)) return rc;
  std::vector<void*> sc(static_cast<size_t>(n));
  for (int k = 0; k < n; ++k) sc[size_t(k)] = static_cast<uint8_t*>(x.kb_scratch.ptr) + offs[size_t(k)];
  return knn_launch_frames(x, frames, cfgs, sc.data(), n, s);
}

}  // extern "C"
