// Sparse keypoint MASDA on gfx950: the candidate graph, the max-sum messages
// and the decision of core/src/match.cpp, for one frame or a batch of frames.
//
// Three launches per batch:
//   k_sp_sort   one workgroup per frame: bitonic sort of the right keypoints by
//               (y, index) in registers and LDS, the epipolar-band index
//               (match.cpp:52-57).
//   k_sp_cand   one wave per left keypoint: 64-ary search of the band, scores
//               in lanes, the kept candidates (match.cpp:61-98, 295-332).
//   k_sp_solve  one 1024-thread workgroup per frame: column lists, damped rho /
//               beta half-steps to convergence, beliefs, margins and the
//               decision (match.cpp:112-231, 334-417), or mutual-NN
//               (match.cpp:419-476).
//
// Compiled with -ffp-contract=off: scores, messages and beliefs round exactly
// as the reference's x86 build does. Every reduction is either order
// independent (max, "max over the others", counts) or broken by an explicit
// total order, so results do not depend on thread scheduling:
//   * right keypoints are ordered by (y, index) (the reference: std::sort);
//   * kept candidates beyond max_candidates by (score desc, band position)
//     (the reference: std::partial_sort);
//   * the greedy completion by (belief desc, score desc, edge index) (the
//     reference: std::sort). Sequential greedy in a total order equals the
//     fixed point of "accept every edge that is first in that order among the
//     remaining edges of both its row and its column", which is what runs here
//     in rounds.
// With no exact ties these orders are the reference's own.
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

constexpr int kSpThreads = 1024;      // solver / sort workgroup
constexpr int kSpWaves = kSpThreads / kWave;
constexpr int kCandWaves = 4;         // left keypoints per candidate workgroup
constexpr int kChunk = 8;             // loads issued together in the solver loops
constexpr float kNeg = -1e30f;

enum SpMode { kModeMasda = 0, kModeMutualNN = 1, kModeCandidates = 2 };

// Everything one sparse frame needs on the device.
struct SpDesc {
  const float4* L;          // left keypoints (x, y, response, local_std)
  const uint64_t* dl;
  const float4* R;
  const uint64_t* dr;
  int nl, nr, K, mode;      // K = max(1, max_candidates)
  int lds_dyn;              // dynamic LDS bytes of the solve launch
  float max_dy, min_d, max_d, sy;   // sy = sigma_y > 0 ? sigma_y : 1
  float w_y, w_prior, w_smooth, damp;
  int iterations;
  float lambda, gamma, eps, ratio;
  int p_cell, p_cols, p_rows, s_n;  // p_cols = 0: no prior
  const float* p_disp;
  const float* p_sigma;
  const uint8_t* p_valid;
  const float* s_disp;      // null: no smoothness prior
  const float* s_sigma;
  const uint8_t* s_valid;
  // scratch
  float* ys;                // sorted right y [nr]
  int* yj;                  // sorted right index [nr]
  int* rc;                  // kept candidates per left keypoint [nl]
  int* ej;                  // edge e = i*K + k: right index, score, messages
  float* es;
  float* rho;
  float* beta;
  float* bel;
  int* col_edges;           // edges grouped by right keypoint [nl*K]
  float* margin;            // [nl]
  int* row_pick;            // [nl]
  int* col_pick;            // [nr]
  int* acc2;                // greedy-completion acceptances [min(nl, nr)]
  uint8_t* lt;              // left taken [nl]
  uint8_t* rt;              // right taken [nr]
  // outputs
  de_match* out;
  int* n_out;
  de_match_stats* stats;
};

struct SpBatch {
  const SpDesc* frames;     // [n]
  const int* start;         // [n + 1] prefix of left keypoints
  int n;
};

__device__ __forceinline__ float fmax_ref(float a, float b) { return (a < b) ? b : a; }

// Order-preserving map of a float to uint32 (-0 and +0 map alike, as `<` sees them).
__device__ __forceinline__ uint32_t ordered(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u << 1) == 0u) u = 0u;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// First index of sorted ys[0, n) where pred turns false (pred holds on a
// prefix), found by 64-ary search across the wave.
template <class Pred>
__device__ int wave_partition_point(const float* ys, int n, Pred pred, int lane) {
  int lo = 0, hi = n;
  while (hi - lo > kWave) {
    const int step = (hi - lo + kWave - 1) / kWave;
    const int pos = lo + lane * step;
    const bool t = pos < hi && pred(ys[pos]);
    const int c = __popcll(__ballot(t));
    const int nlo = c > 0 ? lo + (c - 1) * step + 1 : lo;
    const int nhi = (c < kWave && lo + c * step < hi) ? lo + c * step : hi;
    lo = nlo;
    hi = nhi;
  }
  const int pos = lo + lane;
  const bool t = pos < hi && pred(ys[pos]);
  return lo + __popcll(__ballot(t));
}

// s(i, j) (match.cpp:295-332).
__device__ float pair_score(const SpDesc& f, float4 l, float4 r, uint64_t a, uint64_t b,
                            int i) {
  const float half = 24.0f;   // 0.5 * 48 census bits (match.cpp:340)
  const float desc = (half - float(__popcll(a ^ b))) / half;
  const float dy = (l.y - r.y) / f.sy;
  float s = desc - f.w_y * dy * dy;
  if (f.p_cols > 0 && f.p_rows > 0) {
    const int cx = min(f.p_cols - 1, max(0, int(l.x) / f.p_cell));
    const int cy = min(f.p_rows - 1, max(0, int(l.y) / f.p_cell));
    const int k = cy * f.p_cols + cx;
    if (f.p_valid[k]) {
      const float dp = f.p_disp[k], sd = f.p_sigma[k];
      if (sd > 0.f) {
        const float e = ((l.x - r.x) - dp) / sd;
        s -= f.w_prior * e * e;
      }
    }
  }
  if (f.s_valid && i < f.s_n && f.s_valid[i]) {
    const float sg = f.s_sigma[i];
    if (sg > 0.f) {
      const float e = ((l.x - r.x) - f.s_disp[i]) / sg;
      s -= f.w_smooth * e * e;
    }
  }
  return s;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
  for (int o = 1; o < kWave; o <<= 1) {
    const uint64_t w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  return v;
}

// ---- block-wide helpers (kSpThreads threads) ----

__device__ float block_max(float v, float* sh) {
  for (int o = 1; o < kWave; o <<= 1) v = fmax_ref(v, __shfl_xor(v, o));
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) sh[w] = v;
  __syncthreads();
  float r = sh[0];
  for (int k = 1; k < kSpWaves; ++k) r = fmax_ref(r, sh[k]);
  return r;
}

// Exclusive prefix of v over threads in index order; *total gets the sum.
__device__ int block_excl_scan(int v, int* sh, int* total) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  int x = v;
  for (int o = 1; o < kWave; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  __syncthreads();
  if (lane == kWave - 1) sh[w] = x;
  __syncthreads();
  int before = 0, all = 0;
  for (int k = 0; k < kSpWaves; ++k) {
    if (k < w) before += sh[k];
    all += sh[k];
  }
  *total = all;
  return before + x - v;
}

// Greedy-completion order: belief desc, score desc, edge index asc.
__device__ __forceinline__ bool before_in_greedy(const SpDesc& f, int a, int b) {
  const float ba = f.bel[a], bb = f.bel[b];
  if (ba != bb) return ba > bb;
  const float sa = f.es[a], sb = f.es[b];
  if (sa != sb) return sa > sb;
  return a < b;
}

__device__ void put_match(const SpDesc& f, int slot, int e, float belief) {
  const int i = e / f.K, j = f.ej[e];
  const float4 l = f.L[i], r = f.R[j];
  de_match m;
  m.left = i;
  m.right = j;
  m.score = f.es[e];
  m.belief = belief;
  m.disparity = l.x - r.x;
  m.dy = l.y - r.y;
  m.margin = f.margin[i];
  f.out[slot] = m;
}

// ---------------------------------------------------------------------------
// k_sp_sort: right keypoints by (y, index).

// Up to 2 * kSpThreads keys (every frame of the bench and the tests) the
// network runs in registers: thread t holds positions t and t + kSpThreads,
// partners closer than a wave come by lane shuffles, the pair kSpThreads apart
// is within the thread, and only the distances 64 .. kSpThreads / 2 exchange
// through LDS (two alternating buffers, one barrier a step): 14 barriers at
// 2048 keys instead of 66.
__device__ __forceinline__ void sort_keys_reg(const SpDesc& f, int P, uint64_t* keys) {
  const int t0 = threadIdx.x, t1 = t0 + kSpThreads;
  auto key = [&](int t) -> uint64_t {
    return t < f.nr ? (uint64_t(ordered(f.R[t].y)) << 32) | uint32_t(t) : ~0ull;
  };
  uint64_t a[2] = {key(t0), key(t1)};
  const int tp[2] = {t0, t1};
  int pb = 0;
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j == kSpThreads) {                             // positions t0 < t1 = t0 ^ j
        const bool asc = (t0 & k) == 0;
        const uint64_t lo = a[0] < a[1] ? a[0] : a[1], hi = a[0] < a[1] ? a[1] : a[0];
        a[0] = asc ? lo : hi;
        a[1] = asc ? hi : lo;
        continue;
      }
      uint64_t b[2];
      if (j < kWave) {
        b[0] = __shfl_xor(a[0], j);
        b[1] = __shfl_xor(a[1], j);
      } else {
        uint64_t* buf = keys + pb * (2 * kSpThreads);
        buf[t0] = a[0];
        buf[t1] = a[1];
        __syncthreads();
        b[0] = buf[t0 ^ j];
        b[1] = buf[t1 ^ j];
        pb ^= 1;
      }
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        // ascending blocks put the smaller key at the lower position
        const bool take_min = ((tp[e] & j) == 0) == ((tp[e] & k) == 0);
        const bool lt = a[e] < b[e];
        a[e] = take_min ? (lt ? a[e] : b[e]) : (lt ? b[e] : a[e]);
      }
    }
  }
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int t = tp[e];
    if (t < f.nr) {
      const int jj = int(uint32_t(a[e]));
      f.yj[t] = jj;
      f.ys[t] = f.R[jj].y;
    }
  }
}

__global__ __launch_bounds__(kSpThreads) void k_sp_sort(const SpDesc* __restrict__ frames) {
  extern __shared__ uint64_t keys[];
  const SpDesc f = frames[blockIdx.x];
  int P = 1;
  while (P < f.nr) P <<= 1;
  if (P <= 2 * kSpThreads) {
    sort_keys_reg(f, P, keys);
    return;
  }
  for (int t = threadIdx.x; t < P; t += kSpThreads)
    keys[t] = t < f.nr ? (uint64_t(ordered(f.R[t].y)) << 32) | uint32_t(t) : ~0ull;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < P; t += kSpThreads) {
        const int u = t ^ j;
        if (u > t) {
          const uint64_t a = keys[t], b = keys[u];
          if ((a > b) == ((t & k) == 0)) {
            keys[t] = b;
            keys[u] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int t = threadIdx.x; t < f.nr; t += kSpThreads) {
    const int j = int(uint32_t(keys[t]));
    f.yj[t] = j;
    f.ys[t] = f.R[j].y;
  }
}

// ---------------------------------------------------------------------------
// k_sp_cand: one wave per left keypoint.

__global__ __launch_bounds__(kCandWaves * kWave) void k_sp_cand(SpBatch b, int total) {
  const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * kCandWaves +
                                                  int(threadIdx.x / kWave));
  const int lane = threadIdx.x & (kWave - 1);
  if (item >= total) return;
  int lo = 0, hi = b.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b.start[mid] <= item) lo = mid; else hi = mid - 1;
  }
  const SpDesc& f = b.frames[lo];
  const int i = item - b.start[lo];
  const int K = f.K;
  const float4 l = f.L[i];
  const uint64_t dli = f.dl[i];
  const float ylo = l.y - f.max_dy, yhi = l.y + f.max_dy;
  const int b0 = wave_partition_point(f.ys, f.nr, [&](float y) { return y < ylo; }, lane);
  const int b1 = wave_partition_point(f.ys, f.nr, [&](float y) { return !(yhi < y); }, lane);

  // One band element per lane per chunk of 64 positions.
  auto eval = [&](int p, int* j, float* s) -> bool {
    if (p >= b1) return false;
    *j = f.yj[p];
    const float4 r = f.R[*j];
    const float d = l.x - r.x;
    if (d < f.min_d || d > f.max_d) return false;
    *s = pair_score(f, l, r, dli, f.dr[*j], i);
    return true;
  };
  const bool single = b1 - b0 <= kWave;
  int j0 = -1;
  float s0 = 0.f;
  bool ok0 = false;
  int cnt = 0;
  for (int p0 = b0; p0 < b1; p0 += kWave) {
    int j = -1;
    float s = 0.f;
    const bool ok = eval(p0 + lane, &j, &s);
    cnt += __popcll(__ballot(ok));
    if (single) { ok0 = ok; j0 = j; s0 = s; }
  }
  const int keep = min(cnt, K);
  if (lane == 0) f.rc[i] = keep;
  if (cnt == 0) return;
  const int base = i * K;
  if (cnt <= K) {
    // all of them, in band order
    int w = 0;
    for (int p0 = b0; p0 < b1; p0 += kWave) {
      int j = j0;
      float s = s0;
      bool ok = ok0;
      if (!single) ok = eval(p0 + lane, &j, &s);
      const uint64_t m = __ballot(ok);
      if (ok) {
        const int slot = w + __popcll(m & ((1ull << lane) - 1ull));
        f.ej[base + slot] = j;
        f.es[base + slot] = s;
      }
      w += __popcll(m);
    }
    return;
  }
  // the best K by (score desc, band position asc), in that order
  uint64_t thr = ~0ull;
  for (int r = 0; r < K; ++r) {
    uint64_t bk = 0;
    float bs = 0.f;
    int bj = -1;
    for (int p0 = b0; p0 < b1; p0 += kWave) {
      int j = j0;
      float s = s0;
      bool ok = ok0;
      if (!single) ok = eval(p0 + lane, &j, &s);
      if (ok) {
        const uint64_t key = (uint64_t(ordered(s)) << 32) | uint64_t(0xFFFFFFFFu - uint32_t(p0 + lane));
        if (key < thr && key > bk) { bk = key; bs = s; bj = j; }
      }
    }
    const uint64_t m = wave_max_u64(bk);
    if (bk == m) {
      f.ej[base + r] = bj;
      f.es[base + r] = bs;
    }
    thr = m;
  }
}

// ---------------------------------------------------------------------------
// The decision of match.cpp:409-476 over LDS-resident edges (the solver's
// compacted row-major order c, which orders edges as their index e = i*K + k
// does): beliefs bel[c], scores s[c], right index ej[c], column lists colq,
// row starts rowst. An edge is dead once its row or its column is taken, so
// one flag per edge serves both passes' "still free" tests. The rules and
// orders are those of the global form below (pass 1: row and column agree,
// output in left-keypoint order; pass 2: rounds of locally first edges in
// (belief desc, score desc, edge asc) order, output in that order). Returns
// the number of matches written.
__device__ int decide_lds(const SpDesc& f, de_match_stats& st, const float* bel, const float* sc,
                          const uint16_t* colq, const uint16_t* rowst, const uint16_t* ej,
                          uint8_t* dead, uint16_t* rpick, uint16_t* cpick, uint16_t* acc,
                          const int* col_end, int* shi, int* sh_any, int* sh_n2) {
  constexpr uint16_t kNone = 0xFFFF;
  const int tid = threadIdx.x, nl = f.nl, nr = f.nr, K = f.K;
  auto col_begin = [&](int j) { return j > 0 ? col_end[j - 1] : 0; };
  auto edge_of = [&](int i, int c) { return i * K + (c - int(rowst[i])); };
  auto row_of = [&](int c) {   // the largest i with rowst[i] <= c: then c < rowst[i + 1]
    int lo = 0, hi = nl - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (int(rowst[mid]) <= c) lo = mid; else hi = mid - 1;
    }
    return lo;
  };
  auto before = [&](int a, int b) {
    if (bel[a] != bel[b]) return bel[a] > bel[b];
    if (sc[a] != sc[b]) return sc[a] > sc[b];
    return a < b;
  };
  auto kill_row = [&](int i) {
    for (int c = rowst[i]; c < rowst[i + 1]; ++c) dead[c] = 1;
  };
  auto kill_col = [&](int j) {
    for (int q = col_begin(j); q < col_end[j]; ++q) dead[colq[q]] = 1;
  };
  // best eligible edge per row (first among equals) and per column (smaller
  // edge among equals)
  for (int i = tid; i < nl; i += kSpThreads) {
    int pick = -1;
    for (int c = rowst[i]; c < rowst[i + 1]; ++c) {
      if (sc[c] <= f.lambda) continue;
      if (pick < 0 || bel[c] > bel[pick]) pick = c;
    }
    rpick[i] = pick < 0 ? kNone : uint16_t(pick);
  }
  for (int j = tid; j < nr; j += kSpThreads) {
    int pick = -1;
    for (int q = col_begin(j); q < col_end[j]; ++q) {
      const int c = colq[q];
      if (sc[c] <= f.lambda) continue;
      if (pick < 0 || bel[c] > bel[pick] || (bel[c] == bel[pick] && c < pick)) pick = c;
    }
    cpick[j] = pick < 0 ? kNone : uint16_t(pick);
  }
  __syncthreads();
  // pass 1
  int n1 = 0;
  {
    const int ch = (nl + kSpThreads - 1) / kSpThreads;
    const int r0 = min(nl, tid * ch), r1 = min(nl, r0 + ch);
    auto agreed = [&](int i) { const int c = rpick[i]; return c != kNone && cpick[ej[c]] == c; };
    int cnt = 0;
    for (int i = r0; i < r1; ++i) cnt += agreed(i) ? 1 : 0;
    int pos = block_excl_scan(cnt, shi, &n1);
    for (int i = r0; i < r1; ++i)
      if (agreed(i)) {
        const int c = rpick[i];
        put_match(f, pos++, edge_of(i, c), bel[c]);
      }
    __syncthreads();
    for (int i = r0; i < r1; ++i)
      if (agreed(i)) {
        kill_row(i);
        kill_col(ej[rpick[i]]);
      }
  }
  if (tid == 0) *sh_n2 = 0;
  __syncthreads();
  // pass 2
  for (;;) {
    for (int i = tid; i < nl; i += kSpThreads) {
      int pick = -1;
      for (int c = rowst[i]; c < rowst[i + 1]; ++c) {
        if (dead[c] || sc[c] <= f.lambda) continue;
        if (pick < 0 || before(c, pick)) pick = c;
      }
      rpick[i] = pick < 0 ? kNone : uint16_t(pick);
    }
    for (int j = tid; j < nr; j += kSpThreads) {
      int pick = -1;
      for (int q = col_begin(j); q < col_end[j]; ++q) {
        const int c = colq[q];
        if (dead[c] || sc[c] <= f.lambda) continue;
        if (pick < 0 || before(c, pick)) pick = c;
      }
      cpick[j] = pick < 0 ? kNone : uint16_t(pick);
    }
    if (tid == 0) *sh_any = 0;
    __syncthreads();
    for (int i = tid; i < nl; i += kSpThreads) {
      const int c = rpick[i];
      if (c != kNone && cpick[ej[c]] == c) {
        acc[atomicAdd(sh_n2, 1)] = uint16_t(c);
        *sh_any = 1;
      }
    }
    __syncthreads();
    // the accepted edges' rows and columns leave (after every test of the round)
    for (int i = tid; i < nl; i += kSpThreads) {
      const int c = rpick[i];
      if (c != kNone && cpick[ej[c]] == c) {
        kill_row(i);
        kill_col(ej[c]);
      }
    }
    const int any = *sh_any;
    __syncthreads();
    if (!any) break;
  }
  // pass 2's output order: rank among its acceptances
  const int n2 = *sh_n2;
  for (int a = tid; a < n2; a += kSpThreads) {
    const int c = acc[a];
    int rank = 0;
    for (int b = 0; b < n2; ++b) rank += before(acc[b], c) ? 1 : 0;
    put_match(f, n1 + rank, edge_of(row_of(c), c), bel[c]);
  }
  return n1 + n2;
}

// ---------------------------------------------------------------------------
// k_sp_solve: one workgroup per frame.

__global__ __launch_bounds__(kSpThreads) void k_sp_solve(const SpDesc* __restrict__ frames) {
  extern __shared__ int col_end[];       // [nr]: end of each right keypoint's edge list
  __shared__ float shf[kSpWaves];
  __shared__ int shi[kSpWaves];
  __shared__ int sh_any, sh_n2;
  const SpDesc f = frames[blockIdx.x];
  const int tid = threadIdx.x, nl = f.nl, nr = f.nr, K = f.K;

  // 1. edges grouped by right keypoint (order inside a group is irrelevant:
  //    every reduction over a group is order independent or index-broken)
  for (int j = tid; j < nr; j += kSpThreads) {
    col_end[j] = 0;
    f.rt[j] = 0;
  }
  __syncthreads();
  // A left keypoint's edges (i*K .. i*K+n-1 of ej and es) are read kChunk at a
  // time, every load of a group issued before the first use: one memory
  // latency a group, not an edge (most rows have at most kChunk edges).
  int mine = 0;
  for (int i = tid; i < nl; i += kSpThreads) {
    const int n = f.rc[i];
    mine += n;
    f.lt[i] = 0;
    for (int k0 = 0; k0 < n; k0 += kChunk) {
      int jv[kChunk];
#pragma unroll
      for (int u = 0; u < kChunk; ++u) jv[u] = k0 + u < n ? f.ej[i * K + k0 + u] : 0;
#pragma unroll
      for (int u = 0; u < kChunk; ++u)
        if (k0 + u < n) atomicAdd(&col_end[jv[u]], 1);
    }
  }
  int E = 0;
  (void)block_excl_scan(mine, shi, &E);
  {
    // inclusive prefix of the counts, one contiguous run of columns per thread
    const int ch = (nr + kSpThreads - 1) / kSpThreads;
    const int c0 = min(nr, tid * ch), c1 = min(nr, c0 + ch);
    int run = 0;
    for (int j = c0; j < c1; ++j) run += col_end[j];
    int tot;
    int acc = block_excl_scan(run, shi, &tot);
    __syncthreads();
    for (int j = c0; j < c1; ++j) {
      const int c = col_end[j];
      col_end[j] = acc;          // start, advanced to the end by the fill below
      acc += c;
    }
  }
  __syncthreads();
  for (int i = tid; i < nl; i += kSpThreads) {
    const int n = f.rc[i];
    for (int k0 = 0; k0 < n; k0 += kChunk) {
      int jv[kChunk];
#pragma unroll
      for (int u = 0; u < kChunk; ++u) jv[u] = k0 + u < n ? f.ej[i * K + k0 + u] : 0;
#pragma unroll
      for (int u = 0; u < kChunk; ++u)
        if (k0 + u < n) f.col_edges[atomicAdd(&col_end[jv[u]], 1)] = i * K + k0 + u;
    }
  }
  __syncthreads();
  auto col_begin = [&](int j) { return j > 0 ? col_end[j - 1] : 0; };

  // margins (match.cpp:112-132)
  for (int i = tid; i < nl; i += kSpThreads) {
    const int n = f.rc[i];
    float t = kNeg, r = kNeg;
    for (int k0 = 0; k0 < n; k0 += kChunk) {
      float sv[kChunk];
#pragma unroll
      for (int u = 0; u < kChunk; ++u) sv[u] = k0 + u < n ? f.es[i * K + k0 + u] : kNeg;
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        if (k0 + u >= n) break;
        const float s = sv[u];
        if (s > t) { r = t; t = s; }
        else if (s > r) r = s;
      }
    }
    float m = 0.f;
    if (t != kNeg) m = t - (r == kNeg ? f.lambda : fmax_ref(r, f.lambda));
    f.margin[i] = m;
  }

  de_match_stats st;
  st.left = nl;
  st.right = nr;
  st.candidates = E;
  st.iterations_run = 0;
  st.converged = f.mode == kModeMutualNN ? 1 : 0;
  st.final_max_delta = 0.f;
  st.oscillating = 0;
  if (E == 0) {
    if (tid == 0) {
      *f.n_out = 0;
      *f.stats = st;
    }
    return;
  }

  int n1 = 0;
  if (f.mode == kModeMutualNN) {
    // best edge per right keypoint by score, first in edge order among equals
    for (int j = tid; j < nr; j += kSpThreads) {
      int pick = -1;
      for (int q = col_begin(j); q < col_end[j]; ++q) {
        const int e = f.col_edges[q];
        const float s = f.es[e];
        if (pick < 0 || s > f.es[pick] || (s == f.es[pick] && e < pick)) pick = e;
      }
      f.col_pick[j] = pick;
    }
    __syncthreads();
    const int ch = (nl + kSpThreads - 1) / kSpThreads;
    const int r0 = min(nl, tid * ch), r1 = min(nl, r0 + ch);
    auto accept_nn = [&](int i, int* pick) -> bool {
      const int n = f.rc[i];
      int b1 = -1, b2 = -1;
      for (int k = 0; k < n; ++k) {
        const int e = i * K + k;
        if (b1 < 0 || f.es[e] > f.es[b1]) { b2 = b1; b1 = e; }
        else if (b2 < 0 || f.es[e] > f.es[b2]) b2 = e;
      }
      *pick = b1;
      if (b1 < 0 || f.es[b1] <= f.lambda || f.col_pick[f.ej[b1]] != b1) return false;
      if (b2 >= 0) {
        const float bb = f.es[b1] - f.lambda, s2 = f.es[b2] - f.lambda;
        if (bb > 0.f && s2 > 0.f && s2 / bb > f.ratio) return false;
      }
      return true;
    };
    int cnt = 0, pick;
    for (int i = r0; i < r1; ++i) cnt += accept_nn(i, &pick) ? 1 : 0;
    int pos = block_excl_scan(cnt, shi, &n1);
    for (int i = r0; i < r1; ++i)
      if (accept_nn(i, &pick)) put_match(f, pos++, pick, f.es[pick]);
    if (tid == 0) {
      *f.n_out = n1;
      *f.stats = st;
    }
    return;
  }

  // 2. message passing (match.cpp:362-407)
  float last = 1e30f;
  // When the edge set fits, the messages live in LDS for the whole loop: rho,
  // beta and s per edge in row-major compacted order, the column order as
  // 16-bit compacted indices. The global form is bound by one CU's cache-line
  // request rate (every column gather is a separate line); the arithmetic and
  // its order are the same in both forms.
  const size_t col_bytes = size_t(nr) * 4;
  const size_t need = col_bytes + size_t(E) * 14 + size_t(nl + 1) * 2;
  if (E < 65536 && need <= size_t(f.lds_dyn)) {
    uint8_t* dyn = reinterpret_cast<uint8_t*>(col_end);
    float* rho_s = reinterpret_cast<float*>(dyn + col_bytes);
    float* beta_s = rho_s + E;
    float* es_s = beta_s + E;
    uint16_t* colq = reinterpret_cast<uint16_t*>(es_s + E);
    uint16_t* rowst = colq + E;   // [nl + 1]
    {
      const int ch = (nl + kSpThreads - 1) / kSpThreads;
      const int r0 = min(nl, tid * ch), r1 = min(nl, r0 + ch);
      int run = 0;
      for (int i = r0; i < r1; ++i) run += f.rc[i];
      int tot;
      int acc = block_excl_scan(run, shi, &tot);
      for (int i = r0; i < r1; ++i) {
        rowst[i] = uint16_t(acc);
        acc += f.rc[i];
      }
      if (tid == 0) rowst[nl] = uint16_t(tot);
    }
    __syncthreads();
    for (int i = tid; i < nl; i += kSpThreads) {
      const int c0 = rowst[i], n = rowst[i + 1] - c0;
      for (int k0 = 0; k0 < n; k0 += kChunk) {
        float sv[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) sv[u] = k0 + u < n ? f.es[i * K + k0 + u] : 0.f;
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
          if (k0 + u >= n) break;
          rho_s[c0 + k0 + u] = 0.f;
          beta_s[c0 + k0 + u] = 0.f;
          es_s[c0 + k0 + u] = sv[u];
        }
      }
    }
    for (int q0 = tid; q0 < E; q0 += kChunk * kSpThreads) {
      int ev[kChunk];
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        const int q = q0 + u * kSpThreads;
        ev[u] = q < E ? f.col_edges[q] : 0;
      }
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        const int q = q0 + u * kSpThreads;
        if (q < E) {
          const int i = ev[u] / K;
          colq[q] = uint16_t(rowst[i] + (ev[u] - i * K));
        }
      }
    }
    __syncthreads();
    for (int it = 0; it < f.iterations; ++it) {
      float delta = 0.f;
      for (int i = tid; i < nl; i += kSpThreads) {
        const int c0 = rowst[i], n = rowst[i + 1] - c0;
        float b1 = kNeg, b2 = kNeg;
        int arg = -1;
        for (int k = 0; k < n; ++k) {
          const float v = beta_s[c0 + k];
          if (v > b1) { b2 = b1; b1 = v; arg = k; }
          else if (v > b2) b2 = v;
        }
        for (int k = 0; k < n; ++k) {
          const float other = k == arg ? b2 : b1;
          const float target = es_s[c0 + k] - fmax_ref(f.lambda, other);
          const float ov = rho_s[c0 + k];
          const float nv = (1.f - f.damp) * target + f.damp * ov;
          delta = fmax_ref(delta, fabsf(nv - ov));
          rho_s[c0 + k] = nv;
        }
      }
      __syncthreads();
      for (int j = tid; j < nr; j += kSpThreads) {
        const int q0 = col_begin(j), q1 = col_end[j];
        float b1 = kNeg, b2 = kNeg;
        int arg = -1;
        for (int q = q0; q < q1; ++q) {
          const int c = colq[q];
          const float v = rho_s[c];
          if (v > b1) { b2 = b1; b1 = v; arg = c; }
          else if (v > b2) b2 = v;
        }
        for (int q = q0; q < q1; ++q) {
          const int c = colq[q];
          const float other = c == arg ? b2 : b1;
          const float target = es_s[c] - fmax_ref(f.gamma, other);
          const float ov = beta_s[c];
          const float nv = (1.f - f.damp) * target + f.damp * ov;
          delta = fmax_ref(delta, fabsf(nv - ov));
          beta_s[c] = nv;
        }
      }
      delta = block_max(delta, shf);   // its barriers also order the phases
      if (delta > last) ++st.oscillating;
      last = delta;
      st.iterations_run = it + 1;
      st.final_max_delta = delta;
      if (delta < f.eps) {
        st.converged = 1;
        break;
      }
    }
    __syncthreads();
    const size_t dec_need = need + size_t(nl) * 2 + size_t(nr) * 2 + size_t(min(nl, nr)) * 2;
    if (dec_need <= size_t(f.lds_dyn)) {
      // the decision in LDS too: beliefs over rho's slots, the edges' right
      // index and taken flags over beta's, picks and acceptances after rowst
      for (int c = tid; c < E; c += kSpThreads) rho_s[c] = beta_s[c] + rho_s[c] - es_s[c];
      __syncthreads();
      uint16_t* ej_s = reinterpret_cast<uint16_t*>(beta_s);
      uint8_t* dead = reinterpret_cast<uint8_t*>(ej_s + E);
      for (int i = tid; i < nl; i += kSpThreads) {
        const int c0 = rowst[i], n = rowst[i + 1] - c0;
        for (int k = 0; k < n; ++k) {
          ej_s[c0 + k] = uint16_t(f.ej[i * K + k]);
          dead[c0 + k] = 0;
        }
      }
      uint16_t* rpick = rowst + nl + 1;
      uint16_t* cpick = rpick + nl;
      uint16_t* acc = cpick + nr;
      __syncthreads();
      const int n = decide_lds(f, st, rho_s, es_s, colq, rowst, ej_s, dead, rpick, cpick, acc,
                               col_end, shi, &sh_any, &sh_n2);
      if (tid == 0) {
        *f.n_out = n;
        *f.stats = st;
      }
      return;
    }
    for (int i = tid; i < nl; i += kSpThreads) {
      const int c0 = rowst[i], n = rowst[i + 1] - c0;
      for (int k = 0; k < n; ++k) {
        f.rho[i * K + k] = rho_s[c0 + k];
        f.beta[i * K + k] = beta_s[c0 + k];
      }
    }
    __syncthreads();
    // col_end's region was not touched; the decision below reads it
  } else {
    for (int i = tid; i < nl; i += kSpThreads) {
      const int n = f.rc[i];
      for (int k = 0; k < n; ++k) {
        f.rho[i * K + k] = 0.f;
        f.beta[i * K + k] = 0.f;
      }
    }
    __syncthreads();
    for (int it = 0; it < f.iterations; ++it) {
      float delta = 0.f;
      // rho_ij = s_ij - max(lambda, max_{k != j} beta_ik): competition in a row.
      // Loads are issued kChunk at a time so their latencies overlap; the
      // arithmetic runs in ascending k as the reference's loop does.
      for (int i = tid; i < nl; i += kSpThreads) {
        const int n = f.rc[i], base = i * K;
        float b1 = kNeg, b2 = kNeg;
        int arg = -1;
        for (int k0 = 0; k0 < n; k0 += kChunk) {
          float v[kChunk];
  #pragma unroll
          for (int u = 0; u < kChunk; ++u) v[u] = k0 + u < n ? f.beta[base + k0 + u] : kNeg;
  #pragma unroll
          for (int u = 0; u < kChunk; ++u) {
            if (k0 + u >= n) break;
            if (v[u] > b1) { b2 = b1; b1 = v[u]; arg = k0 + u; }
            else if (v[u] > b2) b2 = v[u];
          }
        }
        for (int k0 = 0; k0 < n; k0 += kChunk) {
          float sv[kChunk], ov[kChunk];
  #pragma unroll
          for (int u = 0; u < kChunk; ++u) {
            sv[u] = k0 + u < n ? f.es[base + k0 + u] : 0.f;
            ov[u] = k0 + u < n ? f.rho[base + k0 + u] : 0.f;
          }
  #pragma unroll
          for (int u = 0; u < kChunk; ++u) {
            if (k0 + u >= n) break;
            const float other = k0 + u == arg ? b2 : b1;
            const float target = sv[u] - fmax_ref(f.lambda, other);
            const float nv = (1.f - f.damp) * target + f.damp * ov[u];
            delta = fmax_ref(delta, fabsf(nv - ov[u]));
            f.rho[base + k0 + u] = nv;
          }
        }
      }
      __syncthreads();
      // beta_ij = s_ij - max(gamma, max_{k != i} rho_kj): competition in a column
      for (int j = tid; j < nr; j += kSpThreads) {
        const int q0 = col_begin(j), q1 = col_end[j];
        float b1 = kNeg, b2 = kNeg;
        int arg = -1;
        for (int c0 = q0; c0 < q1; c0 += kChunk) {
          int ev[kChunk];
          float v[kChunk];
  #pragma unroll
          for (int u = 0; u < kChunk; ++u) ev[u] = c0 + u < q1 ? f.col_edges[c0 + u] : 0;
  #pragma unroll
          for (int u = 0; u < kChunk; ++u) v[u] = c0 + u < q1 ? f.rho[ev[u]] : kNeg;
  #pragma unroll
          for (int u = 0; u < kChunk; ++u) {
            if (c0 + u >= q1) break;
            if (v[u] > b1) { b2 = b1; b1 = v[u]; arg = ev[u]; }
            else if (v[u] > b2) b2 = v[u];
          }
        }
        for (int c0 = q0; c0 < q1; c0 += kChunk) {
          int ev[kChunk];
          float sv[kChunk], ov[kChunk];
  #pragma unroll
          for (int u = 0; u < kChunk; ++u) ev[u] = c0 + u < q1 ? f.col_edges[c0 + u] : 0;
  #pragma unroll
          for (int u = 0; u < kChunk; ++u) {
            sv[u] = c0 + u < q1 ? f.es[ev[u]] : 0.f;
            ov[u] = c0 + u < q1 ? f.beta[ev[u]] : 0.f;
          }
  #pragma unroll
          for (int u = 0; u < kChunk; ++u) {
            if (c0 + u >= q1) break;
            const float other = ev[u] == arg ? b2 : b1;
            const float target = sv[u] - fmax_ref(f.gamma, other);
            const float nv = (1.f - f.damp) * target + f.damp * ov[u];
            delta = fmax_ref(delta, fabsf(nv - ov[u]));
            f.beta[ev[u]] = nv;
          }
        }
      }
      delta = block_max(delta, shf);
      if (delta > last) ++st.oscillating;
      last = delta;
      st.iterations_run = it + 1;
      st.final_max_delta = delta;
      if (delta < f.eps) {
        st.converged = 1;
        break;
      }
    }
  }
  __syncthreads();

  // 3. beliefs; best eligible edge per row (first in edge order) and per column
  for (int i = tid; i < nl; i += kSpThreads) {
    const int n = f.rc[i], base = i * K;
    int pick = -1;
    for (int k = 0; k < n; ++k) {
      const int e = base + k;
      const float b = f.beta[e] + f.rho[e] - f.es[e];
      f.bel[e] = b;
      if (f.es[e] <= f.lambda) continue;
      if (pick < 0 || b > f.bel[pick]) pick = e;
    }
    f.row_pick[i] = pick;
  }
  __syncthreads();
  for (int j = tid; j < nr; j += kSpThreads) {
    int pick = -1;
    for (int q = col_begin(j); q < col_end[j]; ++q) {
      const int e = f.col_edges[q];
      if (f.es[e] <= f.lambda) continue;
      const float b = f.bel[e];
      if (pick < 0 || b > f.bel[pick] || (b == f.bel[pick] && e < pick)) pick = e;
    }
    f.col_pick[j] = pick;
  }
  __syncthreads();

  // 4. pass 1: row and column agree, in left-keypoint order
  {
    const int ch = (nl + kSpThreads - 1) / kSpThreads;
    const int r0 = min(nl, tid * ch), r1 = min(nl, r0 + ch);
    int cnt = 0;
    for (int i = r0; i < r1; ++i) {
      const int e = f.row_pick[i];
      cnt += (e >= 0 && f.col_pick[f.ej[e]] == e) ? 1 : 0;
    }
    int pos = block_excl_scan(cnt, shi, &n1);
    for (int i = r0; i < r1; ++i) {
      const int e = f.row_pick[i];
      if (e >= 0 && f.col_pick[f.ej[e]] == e) {
        put_match(f, pos++, e, f.bel[e]);
        f.lt[i] = 1;
        f.rt[f.ej[e]] = 1;
      }
    }
  }
  if (tid == 0) sh_n2 = 0;
  __syncthreads();

  // 5. pass 2: greedy completion, as rounds of locally first edges
  for (;;) {
    for (int i = tid; i < nl; i += kSpThreads) {
      int pick = -1;
      if (!f.lt[i]) {
        const int n = f.rc[i], base = i * K;
        for (int k = 0; k < n; ++k) {
          const int e = base + k;
          if (f.es[e] <= f.lambda || f.rt[f.ej[e]]) continue;
          if (pick < 0 || before_in_greedy(f, e, pick)) pick = e;
        }
      }
      f.row_pick[i] = pick;
    }
    for (int j = tid; j < nr; j += kSpThreads) {
      int pick = -1;
      if (!f.rt[j]) {
        for (int q = col_begin(j); q < col_end[j]; ++q) {
          const int e = f.col_edges[q];
          if (f.es[e] <= f.lambda || f.lt[e / K]) continue;
          if (pick < 0 || before_in_greedy(f, e, pick)) pick = e;
        }
      }
      f.col_pick[j] = pick;
    }
    if (tid == 0) sh_any = 0;
    __syncthreads();
    for (int i = tid; i < nl; i += kSpThreads) {
      const int e = f.row_pick[i];
      if (e >= 0 && f.col_pick[f.ej[e]] == e) {
        f.acc2[atomicAdd(&sh_n2, 1)] = e;
        f.lt[i] = 1;
        f.rt[f.ej[e]] = 1;
        sh_any = 1;
      }
    }
    __syncthreads();
    const int any = sh_any;
    __syncthreads();
    if (!any) break;
  }
  // Output order of the completion: rank among the accepted edges, counted
  // against tiles of their keys staged in LDS.
  const int n2 = sh_n2;
  // ranking tiles after the column ends (the LDS form's arrays are dead here)
  uint8_t* const tiles = reinterpret_cast<uint8_t*>(col_end) + ((size_t(nr) * 4 + 15) & ~size_t(15));
  int* const tile_e = reinterpret_cast<int*>(tiles);
  float* const tile_b = reinterpret_cast<float*>(tile_e + kSpThreads);
  float* const tile_s = tile_b + kSpThreads;
  for (int a0 = 0; a0 < n2; a0 += kSpThreads) {
    const int a = a0 + tid;
    const int e = a < n2 ? f.acc2[a] : 0;
    const float be = a < n2 ? f.bel[e] : 0.f, se = a < n2 ? f.es[e] : 0.f;
    int rank = 0;
    for (int t0 = 0; t0 < n2; t0 += kSpThreads) {
      __syncthreads();
      if (t0 + tid < n2) {
        const int x = f.acc2[t0 + tid];
        tile_e[tid] = x;
        tile_b[tid] = f.bel[x];
        tile_s[tid] = f.es[x];
      }
      __syncthreads();
      const int cnt = min(kSpThreads, n2 - t0);
      if (a < n2) {
        for (int b = 0; b < cnt; ++b) {
          const float bb = tile_b[b], sb = tile_s[b];
          rank += (bb != be) ? (bb > be) : (sb != se) ? (sb > se) : (tile_e[b] < e);
        }
      }
    }
    if (a < n2) put_match(f, n1 + rank, e, be);
  }
  if (tid == 0) {
    *f.n_out = n1 + n2;
    *f.stats = st;
  }
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Scratch bytes of one frame.
size_t sp_scratch_bytes(int nl, int nr, int K) {
  const size_t E = size_t(nl) * size_t(K);
  const size_t m = size_t(std::min(nl, nr));
  return align256(nr * 4) * 2 + align256(nl * 4) * 4 + align256(E * 4) * 6 +
         align256(nr * 4) + align256(m * 4) + align256(nl) + align256(nr) + 256;
}

void sp_carve(void* base, int nl, int nr, int K, SpDesc* d) {
  uint8_t* p = static_cast<uint8_t*>(base);
  auto take = [&](size_t bytes) { void* r = p; p += align256(bytes); return r; };
  const size_t E = size_t(nl) * size_t(K);
  d->ys = static_cast<float*>(take(size_t(nr) * 4));
  d->yj = static_cast<int*>(take(size_t(nr) * 4));
  d->rc = static_cast<int*>(take(size_t(nl) * 4));
  d->margin = static_cast<float*>(take(size_t(nl) * 4));
  d->row_pick = static_cast<int*>(take(size_t(nl) * 4));
  d->ej = static_cast<int*>(take(E * 4));
  d->es = static_cast<float*>(take(E * 4));
  d->rho = static_cast<float*>(take(E * 4));
  d->beta = static_cast<float*>(take(E * 4));
  d->bel = static_cast<float*>(take(E * 4));
  d->col_edges = static_cast<int*>(take(E * 4));
  d->col_pick = static_cast<int*>(take(size_t(nr) * 4));
  d->acc2 = static_cast<int*>(take(size_t(std::min(nl, nr)) * 4));
  d->lt = static_cast<uint8_t*>(take(size_t(nl)));
  d->rt = static_cast<uint8_t*>(take(size_t(nr)));
  take(size_t(nl) * 4);   // spare, keeps the total equal to sp_scratch_bytes
}

int sp_validate(const de_match_cfg& c, int nl, int nr) {
  if (nl < 0 || nr < 0) return fail(DE_ERR_ARG, "keypoint counts %d, %d", nl, nr);
  if (nr > DE_SPARSE_MAX_RIGHT)
    return fail(DE_ERR_UNSUPPORTED, "%d right keypoints exceed the per-frame limit %d", nr,
                DE_SPARSE_MAX_RIGHT);
  if (c.iterations < 0) return fail(DE_ERR_ARG, "iterations %d < 0", c.iterations);
  if (std::max(1, c.max_candidates) > 1024)
    return fail(DE_ERR_ARG, "max_candidates %d exceeds 1024", c.max_candidates);
  return DE_OK;
}

void sp_fill_cfg(const de_match_cfg& c, int mode, float ratio, SpDesc* d) {
  d->K = std::max(1, c.max_candidates);
  d->mode = mode;
  d->max_dy = c.max_dy;
  d->min_d = c.min_disparity;
  d->max_d = c.max_disparity;
  d->sy = c.sigma_y > 0 ? c.sigma_y : 1.f;
  d->w_y = c.w_y;
  d->w_prior = c.w_prior;
  d->w_smooth = c.w_smooth;
  d->damp = std::min(std::max(c.damping, 0.f), 0.99f);   // match.cpp:362
  d->iterations = c.iterations;
  d->lambda = c.lambda;
  d->gamma = c.gamma;
  d->eps = c.converge_eps;
  d->ratio = ratio;
}

size_t sort_lds_bytes(int nr) {
  int P = 1;
  while (P < nr) P <<= 1;
  return P <= 2 * kSpThreads ? size_t(2) * (2 * kSpThreads) * 8 : size_t(P) * 8;
}

size_t max_sort_lds(const std::vector<SpDesc>& fr) {
  size_t m = 8;
  for (const SpDesc& d : fr) m = std::max(m, sort_lds_bytes(d.nr));
  return m;
}

}  // namespace

// Dynamic LDS of the solve launch: everything the device allows beyond the
// kernel's static arrays (the LDS-resident message loop uses it when the edge
// set fits; the column ends need 4 bytes per right keypoint in any case).
static int solve_dyn_lds(int* bytes) {
  static int cached = -1;
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  if (cached < 0) {
    hipFuncAttributes attr;
    DE_HIP(hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(k_sp_solve)));
    int dev = 0, lds = 0;
    DE_HIP(hipGetDevice(&dev));
    DE_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const int dyn = lds - int(attr.sharedSizeBytes);
    if (dyn < 4 * DE_SPARSE_MAX_RIGHT)
      return fail(DE_ERR_UNSUPPORTED, "sparse solve: %d bytes of LDS left", dyn);
    DE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_sp_solve),
                               hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
    cached = dyn;
  }
  *bytes = cached;
  return DE_OK;
}

// Launch the three stages over frames whose descriptors are at d_frames (device)
// with left-keypoint prefix d_start; `fr` is the host copy.
static int sp_launch(const std::vector<SpDesc>& fr, const SpDesc* d_frames, const int* d_start,
                     int total_left, bool solve, hipStream_t s) {
  const int n = int(fr.size());
  hipLaunchKernelGGL(k_sp_sort, dim3(n), dim3(kSpThreads), max_sort_lds(fr), s, d_frames);
  DE_HIP(hipGetLastError());
  if (total_left > 0) {
    SpBatch b;
    b.frames = d_frames;
    b.start = d_start;
    b.n = n;
    const int blocks = (total_left + kCandWaves - 1) / kCandWaves;
    hipLaunchKernelGGL(k_sp_cand, dim3(blocks), dim3(kCandWaves * kWave), 0, s, b, total_left);
    DE_HIP(hipGetLastError());
  }
  if (solve) {
    int dyn = 0;
    if (int rc = solve_dyn_lds(&dyn)) return rc;
    hipLaunchKernelGGL(k_sp_solve, dim3(n), dim3(kSpThreads), size_t(dyn), s, d_frames);
    DE_HIP(hipGetLastError());
  }
  return DE_OK;
}

// Upload descriptors through the context's descriptor ring and launch.
static int sp_run(Ctx& x, const std::vector<SpDesc>& fr, bool solve, hipStream_t s) {
  const int n = int(fr.size());
  std::vector<int> start(size_t(n) + 1, 0);
  for (int k = 0; k < n; ++k) start[size_t(k) + 1] = start[size_t(k)] + fr[size_t(k)].nl;
  const size_t db = fr.size() * sizeof(SpDesc), sb = start.size() * sizeof(int);
  int dyn = 0;
  if (solve)
    if (int rc = solve_dyn_lds(&dyn)) return rc;
  Ctx::Slot* slot;
  if (int rc = x.acquire_slot(db + sb, &slot)) return rc;
  std::memcpy(slot->host, fr.data(), db);
  for (int k = 0; k < n; ++k) static_cast<SpDesc*>(slot->host)[k].lds_dyn = dyn;
  std::memcpy(static_cast<uint8_t*>(slot->host) + db, start.data(), sb);
  DE_HIP(hipMemcpyAsync(slot->dev, slot->host, db + sb, hipMemcpyHostToDevice, s));
  const SpDesc* d_frames = static_cast<const SpDesc*>(slot->dev);
  const int* d_start = reinterpret_cast<const int*>(static_cast<uint8_t*>(slot->dev) + db);
  if (int rc = sp_launch(fr, d_frames, d_start, start[size_t(n)], solve, s)) return rc;
  DE_HIP(hipEventRecord(slot->done, s));
  slot->used = true;
  return DE_OK;
}

// Host-pointer entry: stage keypoints, descriptors and priors in the context's
// sparse io buffer, run one frame, copy the results back.
static int sp_host(de_ctx* c, const de_keypoint* left, const uint64_t* dl, int nl,
                   const de_keypoint* right, const uint64_t* dr, int nr,
                   const de_match_cfg* cfg, const de_disparity_prior* prior,
                   const de_smooth_prior* smooth, int mode, float ratio, de_match* out,
                   int* n_out, de_match_stats* stats, int* ei, int* ej, float* es) {
  if (!c || !cfg) return fail(DE_ERR_ARG, "sparse: null ctx or cfg");
  if ((nl > 0 && (!left || !dl)) || (nr > 0 && (!right || !dr)))
    return fail(DE_ERR_ARG, "sparse: null keypoints or descriptors");
  if (mode != kModeCandidates && (!n_out || (!out && nl > 0 && nr > 0)))
    return fail(DE_ERR_ARG, "sparse: null output");
  if (int rc = sp_validate(*cfg, nl, nr)) return rc;
  if (prior && prior->cols > 0 && prior->rows > 0 &&
      (prior->cell < 1 || !prior->disparity || !prior->sigma || !prior->valid))
    return fail(DE_ERR_ARG, "sparse: disparity prior needs cell >= 1 and its three arrays");
  if (smooth && smooth->n > 0 && (!smooth->disparity || !smooth->sigma || !smooth->valid))
    return fail(DE_ERR_ARG, "sparse: smoothness prior needs its three arrays");
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  SpDesc d;
  std::memset(&d, 0, sizeof(d));
  sp_fill_cfg(*cfg, mode, ratio, &d);
  d.nl = nl;
  d.nr = nr;
  const int K = d.K;
  const size_t m = size_t(std::max(1, std::min(nl, nr)));
  const int pc = (prior && prior->cols > 0 && prior->rows > 0) ? prior->cols * prior->rows : 0;
  const int sn = (smooth && smooth->n > 0) ? smooth->n : 0;
  // io layout: inputs, priors, outputs
  const size_t in_bytes = align256(size_t(nl) * 16) + align256(size_t(nl) * 8) +
                          align256(size_t(nr) * 16) + align256(size_t(nr) * 8) +
                          align256(size_t(pc) * 9) * 3 + align256(size_t(sn) * 9) * 3;
  const size_t out_bytes = align256(m * sizeof(de_match)) + align256(sizeof(int)) +
                           align256(sizeof(de_match_stats));
  if (int rc = x.sp_io.ensure(in_bytes + out_bytes)) return rc;
  if (int rc = x.sp_scratch.ensure(sp_scratch_bytes(nl, nr, K))) return rc;
  uint8_t* p = static_cast<uint8_t*>(x.sp_io.ptr);
  auto put = [&](const void* h, size_t bytes) -> void* {
    void* r = p;
    if (bytes) (void)hipMemcpyAsync(r, h, bytes, hipMemcpyHostToDevice, x.stream);
    p += align256(bytes);
    return r;
  };
  d.L = static_cast<const float4*>(put(left, size_t(nl) * 16));
  d.dl = static_cast<const uint64_t*>(put(dl, size_t(nl) * 8));
  d.R = static_cast<const float4*>(put(right, size_t(nr) * 16));
  d.dr = static_cast<const uint64_t*>(put(dr, size_t(nr) * 8));
  if (pc > 0) {
    d.p_cell = prior->cell;
    d.p_cols = prior->cols;
    d.p_rows = prior->rows;
    d.p_disp = static_cast<const float*>(put(prior->disparity, size_t(pc) * 4));
    d.p_sigma = static_cast<const float*>(put(prior->sigma, size_t(pc) * 4));
    d.p_valid = static_cast<const uint8_t*>(put(prior->valid, size_t(pc)));
  }
  if (sn > 0) {
    d.s_n = sn;
    d.s_disp = static_cast<const float*>(put(smooth->disparity, size_t(sn) * 4));
    d.s_sigma = static_cast<const float*>(put(smooth->sigma, size_t(sn) * 4));
    d.s_valid = static_cast<const uint8_t*>(put(smooth->valid, size_t(sn)));
  }
  DE_HIP(hipGetLastError());
  p = static_cast<uint8_t*>(x.sp_io.ptr) + in_bytes;
  d.out = reinterpret_cast<de_match*>(p);
  p += align256(m * sizeof(de_match));
  d.n_out = reinterpret_cast<int*>(p);
  p += align256(sizeof(int));
  d.stats = reinterpret_cast<de_match_stats*>(p);
  sp_carve(x.sp_scratch.ptr, nl, nr, K, &d);
  std::vector<SpDesc> fr(1, d);
  if (int rc = sp_run(x, fr, mode != kModeCandidates, x.stream)) return rc;
  if (mode == kModeCandidates) {
    std::vector<int> rc(size_t(std::max(nl, 1))), jj(size_t(nl) * size_t(K) + 1);
    std::vector<float> ss(size_t(nl) * size_t(K) + 1);
    if (nl > 0) {
      DE_HIP(hipMemcpyAsync(rc.data(), d.rc, size_t(nl) * 4, hipMemcpyDeviceToHost, x.stream));
      DE_HIP(hipMemcpyAsync(jj.data(), d.ej, size_t(nl) * K * 4, hipMemcpyDeviceToHost, x.stream));
      DE_HIP(hipMemcpyAsync(ss.data(), d.es, size_t(nl) * K * 4, hipMemcpyDeviceToHost, x.stream));
    }
    DE_HIP(hipStreamSynchronize(x.stream));
    int n = 0;
    for (int i = 0; i < nl; ++i)
      for (int k = 0; k < rc[size_t(i)]; ++k) {
        if (ei) ei[n] = i;
        if (ej) ej[n] = jj[size_t(i) * K + k];
        if (es) es[n] = ss[size_t(i) * K + k];
        ++n;
      }
    if (n_out) *n_out = n;
    return DE_OK;
  }
  int n = 0;
  de_match_stats st;
  DE_HIP(hipMemcpyAsync(&n, d.n_out, sizeof(int), hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipMemcpyAsync(&st, d.stats, sizeof(st), hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipStreamSynchronize(x.stream));
  if (n > 0) {
    DE_HIP(hipMemcpyAsync(out, d.out, size_t(n) * sizeof(de_match), hipMemcpyDeviceToHost,
                          x.stream));
    DE_HIP(hipStreamSynchronize(x.stream));
  }
  *n_out = n;
  if (stats) *stats = st;
  return DE_OK;
}

int sp_validate_cfg(const de_match_cfg& c, int nl, int nr) { return sp_validate(c, nl, nr); }

int sparse_match_dev(Ctx& x, const de_keypoint* L, const uint64_t* dl, int nl,
                     const de_keypoint* R, const uint64_t* dr, int nr, const de_match_cfg& cfg,
                     de_match* out, int* n_out, de_match_stats* stats, hipStream_t s) {
  if (int rc = sp_validate(cfg, nl, nr)) return rc;
  SpDesc d;
  std::memset(&d, 0, sizeof(d));
  sp_fill_cfg(cfg, kModeMasda, 0.f, &d);
  d.L = reinterpret_cast<const float4*>(L);
  d.dl = dl;
  d.R = reinterpret_cast<const float4*>(R);
  d.dr = dr;
  d.nl = nl;
  d.nr = nr;
  d.out = out;
  d.n_out = n_out;
  d.stats = stats;
  if (int rc = x.sp_scratch.ensure(sp_scratch_bytes(nl, nr, d.K))) return rc;
  sp_carve(x.sp_scratch.ptr, nl, nr, d.K, &d);
  std::vector<SpDesc> fr(1, d);
  return sp_run(x, fr, true, s);
}

}  // namespace de

using namespace de;

extern "C" {

void de_match_default_cfg(de_match_cfg* c) {
  if (!c) return;
  // match.hpp:136-176
  c->max_dy = 2.0f;
  c->min_disparity = 1.0f;
  c->max_disparity = 220.0f;
  c->max_candidates = 24;
  c->sigma_y = 1.0f;
  c->w_y = 1.0f;
  c->w_prior = 1.0f;
  c->w_smooth = 1.0f;
  c->iterations = 20;
  c->damping = 0.4f;
  c->lambda = 0.0f;
  c->gamma = 0.0f;
  c->converge_eps = 1e-4f;
}

int de_match_masda(de_ctx* c, const de_keypoint* left, const uint64_t* dl, int nl,
                   const de_keypoint* right, const uint64_t* dr, int nr,
                   const de_match_cfg* cfg, const de_disparity_prior* prior,
                   const de_smooth_prior* smooth, de_match* out, int* n_out,
                   de_match_stats* stats) {
  return sp_host(c, left, dl, nl, right, dr, nr, cfg, prior, smooth, kModeMasda, 0.f, out,
                 n_out, stats, nullptr, nullptr, nullptr);
}

int de_match_mutual_nn(de_ctx* c, const de_keypoint* left, const uint64_t* dl, int nl,
                       const de_keypoint* right, const uint64_t* dr, int nr,
                       const de_match_cfg* cfg, const de_disparity_prior* prior,
                       const de_smooth_prior* smooth, float ratio, de_match* out,
                       int* n_out, de_match_stats* stats) {
  return sp_host(c, left, dl, nl, right, dr, nr, cfg, prior, smooth, kModeMutualNN, ratio,
                 out, n_out, stats, nullptr, nullptr, nullptr);
}

int de_match_candidates(de_ctx* c, const de_keypoint* left, const uint64_t* dl, int nl,
                        const de_keypoint* right, const uint64_t* dr, int nr,
                        const de_match_cfg* cfg, const de_disparity_prior* prior,
                        const de_smooth_prior* smooth, int* ei, int* ej, float* es,
                        int* n_edges) {
  if (!n_edges) return fail(DE_ERR_ARG, "de_match_candidates: null n_edges");
  return sp_host(c, left, dl, nl, right, dr, nr, cfg, prior, smooth, kModeCandidates, 0.f,
                 nullptr, n_edges, nullptr, ei, ej, es);
}

int de_match_masda_batch_dev(de_ctx* c, const de_sparse_frame* frames,
                             const de_match_cfg* cfgs, int n, void* stream) {
  if (!c || !frames || !cfgs || n < 1) return fail(DE_ERR_ARG, "sparse batch: %d frames", n);
  for (int k = 0; k < n; ++k) {
    const de_sparse_frame& q = frames[k];
    if (int rc = sp_validate(cfgs[k], q.n_left, q.n_right)) return rc;
    if ((q.n_left > 0 && (!q.left || !q.desc_left)) ||
        (q.n_right > 0 && (!q.right || !q.desc_right)) || !q.out || !q.n_out || !q.stats)
      return fail(DE_ERR_ARG, "sparse batch frame %d: null pointer", k);
  }
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : x.stream;
  std::vector<SpDesc> fr(static_cast<size_t>(n));
  size_t total = 0;
  std::vector<size_t> off(static_cast<size_t>(n));
  for (int k = 0; k < n; ++k) {
    off[size_t(k)] = total;
    total += align256(sp_scratch_bytes(frames[k].n_left, frames[k].n_right,
                                       std::max(1, cfgs[k].max_candidates)));
  }
  if (int rc = x.sp_scratch.ensure(total)) return rc;
  for (int k = 0; k < n; ++k) {
    const de_sparse_frame& q = frames[k];
    SpDesc& d = fr[size_t(k)];
    std::memset(&d, 0, sizeof(d));
    sp_fill_cfg(cfgs[k], kModeMasda, 0.f, &d);
    d.L = reinterpret_cast<const float4*>(q.left);
    d.dl = q.desc_left;
    d.R = reinterpret_cast<const float4*>(q.right);
    d.dr = q.desc_right;
    d.nl = q.n_left;
    d.nr = q.n_right;
    d.out = q.out;
    d.n_out = q.n_out;
    d.stats = q.stats;
    sp_carve(static_cast<uint8_t*>(x.sp_scratch.ptr) + off[size_t(k)], d.nl, d.nr, d.K, &d);
  }
  return sp_run(x, fr, true, s);
}

}  // extern "C"
