// Edge-list MASDA in float64 on gfx950: the NumPy call shapes of the
// reference's article code (article/masda_stereo.py, article/dense_masda.py).
//
//   dense_problem       dense_masda.py:39-79: every interior pixel against every
//                       candidate on its row; edges d-major, then y, then x;
//                       s = (24 - Hamming) / 24 in float64, stored as float32.
//   _seg_max_excluding  masda_stereo.py:417-442: per-segment max of the other
//                       elements (max, count at max, max below).
//   masda_sparse        masda_stereo.py:453-492: float64 rho/beta half-steps,
//                       damping, convergence on the largest change, belief =
//                       beta + rho - s, greedy decision in belief order.
//   masda / decide      masda_stereo.py:329-414: the same on a dense score
//                       matrix (-inf = no edge), edges in row-major order.
//   census / hamming /  masda_stereo.py:264-326: descriptors at points,
//   build_problem       popcount distances, the keypoint score matrix.
//
// Layout: the edges stay in the caller's order; two counting sorts give
// segment lists by left index (rows) and by right index (columns). Each
// half-step is one thread per segment: a pass for (max, count at max, max
// below), then the damped update of its edges in place. Iterations run from a
// queue of launches; once the largest change drops below eps a device flag
// makes the remaining launches return at once, so nothing waits on the host.
// The decision takes, in rounds, every edge that comes first in (belief desc,
// edge index asc) among the remaining eligible edges of both its row and its
// column: that is the sequential greedy result in that order. The reference
// orders by numpy's unstable argsort, so exactly equal beliefs may be taken in
// another order there. Compiled with -ffp-contract=off: every float64
// operation rounds as numpy's does.
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

constexpr int kT = 256;

__device__ __forceinline__ double np_maximum(double a, double b) {
  // numpy.maximum: NaN propagates; otherwise the larger value
  if (a != a) return a;
  if (b != b) return b;
  return a < b ? b : a;
}

__device__ __forceinline__ unsigned long long dord(double v) {   // order-preserving
  unsigned long long u = __double_as_longlong(v);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double dord_inv(unsigned long long u) {
  return __longlong_as_double((u & 0x8000000000000000ull) ? (u & 0x7FFFFFFFFFFFFFFFull) : ~u);
}

struct EdgeSet {
  const int64_t* ei64;    // caller's indices (int64)
  const int64_t* ej64;
  const double* se;       // scores
  int* ei;                // int32 copies
  int* ej;
  int E, m, n;
  int* row_off;           // [m + 1]
  int* row_edges;         // [E]
  int* col_off;           // [n + 1]
  int* col_edges;         // [E]
  int* cursor;            // [max(m, n)] fill cursors
  double* rho;            // [E]
  double* beta;           // [E]
  double* bel;            // [E]
  unsigned long long* delta_bits;   // [iters]: order-mapped max change per iteration
  double* hist;           // [iters]
  int* state;             // [0] error, [1] converged, [2] iterations run, [3] round progress
  int* rbest;             // [m]
  int* cbest;             // [n]
  uint8_t* used_i;        // [m]
  uint8_t* used_j;        // [n]
  int* assign;            // [m]: right index or -1
};

// -- preparation ----------------------------------------------------------

__global__ void k_ed_validate(EdgeSet s) {
  const int e = blockIdx.x * kT + threadIdx.x;
  if (e >= s.E) return;
  const int64_t i = s.ei64[e], j = s.ej64[e];
  if (i < 0 || i >= s.m || j < 0 || j >= s.n) {
    atomicOr(&s.state[0], 1);
    s.ei[e] = 0;
    s.ej[e] = 0;
    return;
  }
  s.ei[e] = int(i);
  s.ej[e] = int(j);
}

__global__ void k_ed_zero_int(int* p, int n) {
  const int t = blockIdx.x * kT + threadIdx.x;
  if (t < n) p[t] = 0;
}

__global__ void k_ed_hist(EdgeSet s) {
  const int e = blockIdx.x * kT + threadIdx.x;
  if (e >= s.E || s.state[0]) return;
  atomicAdd(&s.row_off[s.ei[e] + 1], 1);
  atomicAdd(&s.col_off[s.ej[e] + 1], 1);
}

// Inclusive prefix over off[1..n] in place, one workgroup (n up to a few
// million: each thread scans a contiguous run).
__global__ __launch_bounds__(1024) void k_ed_scan(int* off, int n) {
  __shared__ int sh[16];
  const int tid = threadIdx.x;
  const int ch = (n + 1023) / 1024;
  const int a = 1 + min(n, tid * ch), b = 1 + min(n, tid * ch + ch);
  int run = 0;
  for (int k = a; k < b; ++k) run += off[k];
  const int lane = tid & 63, w = tid >> 6;
  int x = run;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  int before = 0;
  for (int k = 0; k < w; ++k) before += sh[k];
  int acc = before + x - run;
  for (int k = a; k < b; ++k) {
    acc += off[k];
    off[k] = acc;
  }
  if (tid == 0) off[0] = 0;
}

__global__ void k_ed_copy_int(int* dst, const int* src, int n) {
  const int t = blockIdx.x * kT + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

__global__ void k_ed_fill(EdgeSet s, int by_col) {
  const int e = blockIdx.x * kT + threadIdx.x;
  if (e >= s.E || s.state[0]) return;
  if (by_col) s.col_edges[atomicAdd(&s.cursor[s.ej[e]], 1)] = e;
  else s.row_edges[atomicAdd(&s.cursor[s.ei[e]], 1)] = e;
}

__global__ void k_ed_init(EdgeSet s, int iters) {
  const int e = blockIdx.x * kT + threadIdx.x;
  if (e < s.E) {
    s.rho[e] = 0.0;
    s.beta[e] = 0.0;
  }
  if (e < iters) s.delta_bits[e] = 0ull;
  if (e == 0) {
    s.state[1] = 0;
    s.state[2] = 0;
  }
}

// -- message passing --------------------------------------------------------

// One half-step over segments (rows: in = beta, out = rho, floor = lambda;
// columns: in = rho, out = beta, floor = gamma). The max-excluding-self of an
// element: below the max it is the max; at the max it is the max when another
// element also attains it, else the largest value below the max.
__global__ void k_ed_half(EdgeSet s, int cols, double floor_v, double damping, int it) {
  if (s.state[0] || s.state[1]) return;
  const int g = blockIdx.x * kT + threadIdx.x;
  const int nseg = cols ? s.n : s.m;
  const int* off = cols ? s.col_off : s.row_off;
  const int* lst = cols ? s.col_edges : s.row_edges;
  const double* in = cols ? s.rho : s.beta;
  double* out = cols ? s.beta : s.rho;
  double dmax = 0.0;
  if (g < nseg) {
    const int q0 = off[g], q1 = off[g + 1];
    double m1 = -INFINITY, m2 = -INFINITY;
    int cnt = 0;
    for (int q = q0; q < q1; ++q) {
      const double v = in[lst[q]];
      if (v > m1) m1 = v;
    }
    for (int q = q0; q < q1; ++q) {
      const double v = in[lst[q]];
      if (v >= m1) cnt += 1;
      else if (v > m2) m2 = v;
    }
    const double second = cnt > 1 ? m1 : m2;
    const double keep = 1.0 - damping;
    for (int q = q0; q < q1; ++q) {
      const int e = lst[q];
      const double v = in[e];
      const double excl = v >= m1 ? second : m1;
      const double comp = np_maximum(floor_v, excl);
      const double old = out[e];
      const double nv = keep * (s.se[e] - comp) + damping * old;
      const double d = fabs(nv - old);
      dmax = d > dmax ? d : dmax;
      out[e] = nv;
    }
  }
  // block max of the change, then one atomic per block (values are >= 0)
  for (int o = 1; o < 64; o <<= 1) {
    const double t = __shfl_xor(dmax, o);
    dmax = t > dmax ? t : dmax;
  }
  __shared__ double sh[kT / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = dmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sh[0];
    for (int k = 1; k < kT / 64; ++k) b = sh[k] > b ? sh[k] : b;
    atomicMax(&s.delta_bits[it], dord(b));
  }
}

// _seg_max_excluding over the column segments: out[e] = max of the other
// elements of e's segment (-inf when alone).
__global__ void k_ed_segmax(EdgeSet s, const double* vals, double* out) {
  const int g = blockIdx.x * kT + threadIdx.x;
  if (g >= s.n || s.state[0]) return;
  const int q0 = s.col_off[g], q1 = s.col_off[g + 1];
  double m1 = -INFINITY, m2 = -INFINITY;
  int cnt = 0;
  for (int q = q0; q < q1; ++q) {
    const double v = vals[s.col_edges[q]];
    if (v > m1) m1 = v;
  }
  for (int q = q0; q < q1; ++q) {
    const double v = vals[s.col_edges[q]];
    if (v >= m1) cnt += 1;
    else if (v > m2) m2 = v;
  }
  const double second = cnt > 1 ? m1 : m2;
  for (int q = q0; q < q1; ++q) {
    const int e = s.col_edges[q];
    out[e] = vals[e] >= m1 ? second : m1;
  }
}

__global__ void k_ed_close_iter(EdgeSet s, int it, double eps) {
  if (s.state[0] || s.state[1]) return;
  const double d = dord_inv(s.delta_bits[it]);
  s.hist[it] = d;
  s.state[2] = it + 1;
  if (d < eps) s.state[1] = 1;
}

__global__ void k_ed_belief(EdgeSet s) {
  const int e = blockIdx.x * kT + threadIdx.x;
  if (e >= s.E || s.state[0]) return;
  s.bel[e] = s.beta[e] + s.rho[e] - s.se[e];
}

// -- decision ----------------------------------------------------------------

__global__ void k_ed_dec_init(EdgeSet s) {
  const int t = blockIdx.x * kT + threadIdx.x;
  if (t < s.m) {
    s.used_i[t] = 0;
    s.assign[t] = -1;
  }
  if (t < s.n) s.used_j[t] = 0;
}

// a before b: belief desc, then edge index asc
__device__ __forceinline__ bool ed_before(const EdgeSet& s, int a, int b) {
  const double ba = s.bel[a], bb = s.bel[b];
  if (ba != bb) return ba > bb;
  return a < b;
}

__device__ __forceinline__ bool ed_eligible(const EdgeSet& s, int e, double lam) {
  const double b = s.bel[e];
  return s.se[e] > lam && b == b && b != -INFINITY;
}

__global__ void k_ed_round_best(EdgeSet s, int cols, double lam) {
  if (s.state[0]) return;
  const int g = blockIdx.x * kT + threadIdx.x;
  if (g == 0 && !cols) s.state[3] = 0;
  const int nseg = cols ? s.n : s.m;
  if (g >= nseg) return;
  int* best = cols ? s.cbest : s.rbest;
  if (cols ? s.used_j[g] : s.used_i[g]) {
    best[g] = -1;
    return;
  }
  const int* off = cols ? s.col_off : s.row_off;
  const int* lst = cols ? s.col_edges : s.row_edges;
  int pick = -1;
  for (int q = off[g]; q < off[g + 1]; ++q) {
    const int e = lst[q];
    if (!ed_eligible(s, e, lam)) continue;
    if (cols ? s.used_i[s.ei[e]] : s.used_j[s.ej[e]]) continue;
    if (pick < 0 || ed_before(s, e, pick)) pick = e;
  }
  best[g] = pick;
}

__global__ void k_ed_round_accept(EdgeSet s) {
  if (s.state[0]) return;
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= s.m) return;
  const int e = s.rbest[i];
  if (e < 0) return;
  const int j = s.ej[e];
  if (s.cbest[j] != e) return;
  s.assign[i] = j;
  s.used_i[i] = 1;
  s.used_j[j] = 1;
  s.state[3] = 1;
}

// -- dense problem -------------------------------------------------------------

struct DpDesc {
  const uint64_t* cl;
  const uint64_t* cr;
  int W, H, hw, hh, dmin, nd;
  const long long* start;   // [nd + 1] edge prefix per disparity
  int64_t* ei;
  int64_t* ej;
  float* se;
  int ny;
};

__global__ void k_dense_problem(DpDesc p, long long E) {
  const long long e = (long long)blockIdx.x * kT + threadIdx.x;
  if (e >= E) return;
  int lo = 0, hi = p.nd - 1;   // the disparity block holding e
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (p.start[mid] <= e) lo = mid; else hi = mid - 1;
  }
  const int d = p.dmin + lo;
  const long long k = e - p.start[lo];
  const int nx = p.W - 2 * p.hw - d;
  const int y = p.hh + int(k / nx);
  const int x = p.hw + d + int(k % nx);
  const size_t li = size_t(y) * p.W + x;
  const int h = __popcll(p.cl[li] ^ p.cr[li - d]);
  p.ei[e] = (int64_t)li;
  p.ej[e] = (int64_t)(li - d);
  p.se[e] = float((24.0 - double(h)) / 24.0);
}

// dense matrix -> edges: row-major finite entries
__global__ void k_mat_count(const float* S, int m, int n, int* cnt) {
  const int i = blockIdx.x;
  int c = 0;
  for (int j = threadIdx.x; j < n; j += kT) c += isfinite(S[size_t(i) * n + j]) ? 1 : 0;
  for (int o = 1; o < 64; o <<= 1) c += __shfl_xor(c, o);
  __shared__ int sh[kT / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int k = 0; k < kT / 64; ++k) t += sh[k];
    cnt[i + 1] = t;
  }
}

__global__ void k_mat_edges(const float* S, int m, int n, const int* off, int64_t* ei,
                            int64_t* ej, double* se) {
  // one workgroup per row, ballot-ordered compaction (row-major order kept)
  const int i = blockIdx.x;
  __shared__ int base_sh;
  __shared__ int wsum[kT / 64];
  if (threadIdx.x == 0) base_sh = off[i];
  __syncthreads();
  for (int j0 = 0; j0 < n; j0 += kT) {
    const int j = j0 + threadIdx.x;
    const bool ok = j < n && isfinite(S[size_t(i) * n + j]);
    const unsigned long long bal = __ballot(ok);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wsum[w] = __popcll(bal);
    __syncthreads();
    int before = base_sh;
    for (int k = 0; k < w; ++k) before += wsum[k];
    if (ok) {
      const int pos = before + __popcll(bal & ((1ull << lane) - 1ull));
      ei[pos] = i;
      ej[pos] = j;
      se[pos] = double(S[size_t(i) * n + j]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int k = 0; k < kT / 64; ++k) t += wsum[k];
      base_sh += t;
    }
    __syncthreads();
  }
}

__global__ void k_f32_to_f64(const float* a, double* b, int n) {
  const int t = blockIdx.x * kT + threadIdx.x;
  if (t < n) b[t] = double(a[t]);
}

__global__ void k_assign_disp(const int* assign, int W, int m, float* disp) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= m) return;
  const int j = assign[i];
  disp[i] = j < 0 ? NAN : float((i % W) - (j % W));
}

int gblocks(long long n) { return int((n + kT - 1) / kT); }

size_t a256(size_t v) { return (v + 255) & ~size_t(255); }

// Scratch for an edge set of E edges over m rows and n columns.
size_t edge_scratch_bytes(int E, int m, int n, int iters) {
  const size_t e = size_t(std::max(E, 1)), mm = size_t(std::max(m, 1)), nn = size_t(std::max(n, 1));
  return a256(e * 4) * 4 + a256((mm + 1) * 4) + a256((nn + 1) * 4) + a256(std::max(mm, nn) * 4) +
         a256(e * 8) * 3 + a256(size_t(std::max(iters, 1)) * 8) * 2 + a256(16) + a256(mm * 4) * 2 +
         a256(nn * 4) + a256(mm) + a256(nn);
}

void edge_carve(void* base, int E, int m, int n, int iters, EdgeSet* s) {
  uint8_t* p = static_cast<uint8_t*>(base);
  auto take = [&](size_t b) { void* r = p; p += a256(b); return r; };
  const size_t e = size_t(std::max(E, 1)), mm = size_t(std::max(m, 1)), nn = size_t(std::max(n, 1));
  s->ei = static_cast<int*>(take(e * 4));
  s->ej = static_cast<int*>(take(e * 4));
  s->row_edges = static_cast<int*>(take(e * 4));
  s->col_edges = static_cast<int*>(take(e * 4));
  s->row_off = static_cast<int*>(take((mm + 1) * 4));
  s->col_off = static_cast<int*>(take((nn + 1) * 4));
  s->cursor = static_cast<int*>(take(std::max(mm, nn) * 4));
  s->rho = static_cast<double*>(take(e * 8));
  s->beta = static_cast<double*>(take(e * 8));
  s->bel = static_cast<double*>(take(e * 8));
  s->delta_bits = static_cast<unsigned long long*>(take(size_t(std::max(iters, 1)) * 8));
  s->hist = static_cast<double*>(take(size_t(std::max(iters, 1)) * 8));
  s->state = static_cast<int*>(take(16));
  s->rbest = static_cast<int*>(take(mm * 4));
  s->assign = static_cast<int*>(take(mm * 4));
  s->cbest = static_cast<int*>(take(nn * 4));
  s->used_i = static_cast<uint8_t*>(take(mm));
  s->used_j = static_cast<uint8_t*>(take(nn));
}

}  // namespace

// The solver on an edge set whose ei64/ej64/se and scratch are set. Runs the
// message passing (iters > 0) and/or the decision (decide), synchronising
// only between decision rounds.
static int edge_solve(Ctx& x, EdgeSet& s, int iters, double lam, double gam, double damping,
                      double eps, bool messages, bool decide, hipStream_t st) {
  const int E = s.E;
  int zero4[4] = {0, 0, 0, 0};
  DE_HIP(hipMemcpyAsync(s.state, zero4, 16, hipMemcpyHostToDevice, st));
  if (E > 0) {
    hipLaunchKernelGGL(k_ed_validate, dim3(gblocks(E)), dim3(kT), 0, st, s);
    hipLaunchKernelGGL(k_ed_zero_int, dim3(gblocks(s.m + 1)), dim3(kT), 0, st, s.row_off, s.m + 1);
    hipLaunchKernelGGL(k_ed_zero_int, dim3(gblocks(s.n + 1)), dim3(kT), 0, st, s.col_off, s.n + 1);
    hipLaunchKernelGGL(k_ed_hist, dim3(gblocks(E)), dim3(kT), 0, st, s);
    hipLaunchKernelGGL(k_ed_scan, dim3(1), dim3(1024), 0, st, s.row_off, s.m);
    hipLaunchKernelGGL(k_ed_scan, dim3(1), dim3(1024), 0, st, s.col_off, s.n);
    hipLaunchKernelGGL(k_ed_copy_int, dim3(gblocks(s.m)), dim3(kT), 0, st, s.cursor, s.row_off, s.m);
    hipLaunchKernelGGL(k_ed_fill, dim3(gblocks(E)), dim3(kT), 0, st, s, 0);
    hipLaunchKernelGGL(k_ed_copy_int, dim3(gblocks(s.n)), dim3(kT), 0, st, s.cursor, s.col_off, s.n);
    hipLaunchKernelGGL(k_ed_fill, dim3(gblocks(E)), dim3(kT), 0, st, s, 1);
    DE_HIP(hipGetLastError());
  }
  if (messages) {
    hipLaunchKernelGGL(k_ed_init, dim3(gblocks(std::max(E, iters))), dim3(kT), 0, st, s, iters);
    for (int it = 0; it < iters && E > 0; ++it) {
      hipLaunchKernelGGL(k_ed_half, dim3(gblocks(s.m)), dim3(kT), 0, st, s, 0, lam, damping, it);
      hipLaunchKernelGGL(k_ed_half, dim3(gblocks(s.n)), dim3(kT), 0, st, s, 1, gam, damping, it);
      hipLaunchKernelGGL(k_ed_close_iter, dim3(1), dim3(1), 0, st, s, it, eps);
    }
    if (E > 0) hipLaunchKernelGGL(k_ed_belief, dim3(gblocks(E)), dim3(kT), 0, st, s);
    DE_HIP(hipGetLastError());
  }
  int state[4] = {0, 0, 0, 0};
  if (decide) {
    hipLaunchKernelGGL(k_ed_dec_init, dim3(gblocks(std::max(s.m, s.n))), dim3(kT), 0, st, s);
    for (int round = 0; E > 0; ++round) {
      // a few rounds per host check; a round with no progress leaves the state
      for (int r = 0; r < 4; ++r) {
        hipLaunchKernelGGL(k_ed_round_best, dim3(gblocks(s.m)), dim3(kT), 0, st, s, 0, lam);
        hipLaunchKernelGGL(k_ed_round_best, dim3(gblocks(s.n)), dim3(kT), 0, st, s, 1, lam);
        hipLaunchKernelGGL(k_ed_round_accept, dim3(gblocks(s.m)), dim3(kT), 0, st, s);
      }
      DE_HIP(hipGetLastError());
      DE_HIP(hipMemcpyAsync(state, s.state, 16, hipMemcpyDeviceToHost, st));
      DE_HIP(hipStreamSynchronize(st));
      if (state[0] || !state[3]) break;
      if (round > s.m) return fail(DE_ERR_HIP, "decision did not settle");
    }
  }
  DE_HIP(hipMemcpyAsync(state, s.state, 16, hipMemcpyDeviceToHost, st));
  DE_HIP(hipStreamSynchronize(st));
  if (state[0]) return fail(DE_ERR_ARG, "edge index out of range (ei in [0, m), ej in [0, n))");
  return DE_OK;
}

}  // namespace de

using namespace de;

namespace {

// host-pointer staging: copy arrays to the context's edge io buffer
// ---- census / hamming / build_problem (masda_stereo.py:264-326) ----

// census(img, pts, half_w, half_h): one thread per point; bit b in scan order
// dy-major, dx-minor, centre skipped, set when the neighbour is darker. Indices
// follow numpy's: a negative row or column counts from the end (the host has
// already refused any index numpy would refuse).
__global__ __launch_bounds__(kT) void k_census_points(const uint8_t* __restrict__ img, int W, int H,
                                                      const int64_t* __restrict__ xy, int64_t n,
                                                      int hw, int hh, uint64_t* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (i >= n) return;
  const int64_t x = xy[2 * i], y = xy[2 * i + 1];
  auto at = [&](int64_t yy, int64_t xx) -> int {
    if (yy < 0) yy += H;
    if (xx < 0) xx += W;
    return img[size_t(yy) * W + size_t(xx)];
  };
  const int c = at(y, x);
  uint64_t bits = 0;
  int b = 0;
  for (int dy = -hh; dy <= hh; ++dy)
    for (int dx = -hw; dx <= hw; ++dx) {
      if (dx == 0 && dy == 0) continue;
      if (at(y + dy, x + dx) < c) bits |= uint64_t(1) << b;
      ++b;
    }
  out[i] = bits;
}

// hamming(a, b): popcount of a ^ b, elementwise (broadcast on the host).
__global__ __launch_bounds__(kT) void k_hamming(const uint64_t* __restrict__ a,
                                                const uint64_t* __restrict__ b, int64_t n,
                                                int32_t* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (i < n) out[i] = __popcll(a[i] ^ b[i]);
}

// build_problem: S[i, j] = (BITS/2 - h)/(BITS/2) - w_y (dy/sigma_y)^2 where
// |dy| <= max_dy and dmin <= d <= dmax, -inf elsewhere; one thread per (i, j).
// T is the points' type: with float32 points numpy evaluates dy, d, the gate
// and the penalty in float32 (the Python-float parameters take the array's
// type) and desc in float64; the difference is float64, stored as float32.
// With float64 points everything is float64 until the store.
constexpr double kBits = 48.0;     // masda_stereo.py:302, the 7x7 census length
template <class T>
__global__ __launch_bounds__(kT) void k_build_problem(const T* __restrict__ pl,
                                                      const uint64_t* __restrict__ dl, int m,
                                                      const T* __restrict__ pr,
                                                      const uint64_t* __restrict__ dr, int n,
                                                      T max_dy, T dmin, T dmax, T sigma_y,
                                                      T w_y, float* __restrict__ S) {
  const int64_t e = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (e >= int64_t(m) * n) return;
  const int i = int(e / n), j = int(e - int64_t(i) * n);
  const T dy = pl[2 * i + 1] - pr[2 * j + 1];
  const T d = pl[2 * i] - pr[2 * j];
  const T ady = dy < T(0) ? -dy : dy;
  float v = -INFINITY;
  if (ady <= max_dy && d >= dmin && d <= dmax) {
    const int h = __popcll(dl[i] ^ dr[j]);
    const double desc = (kBits / 2 - double(h)) / (kBits / 2);
    const T q = dy / sigma_y;
    const T pen = w_y * (q * q);
    v = float(desc - double(pen));
  }
  S[e] = v;
}

struct Staging {
  uint8_t* p;
  void* put(const void* h, size_t bytes, hipStream_t s) {
    void* r = p;
    if (bytes && h) (void)hipMemcpyAsync(r, h, bytes, hipMemcpyHostToDevice, s);
    p += a256(bytes);
    return r;
  }
  void* take(size_t bytes) {
    void* r = p;
    p += a256(bytes);
    return r;
  }
};

int run_edges_host(de_ctx* c, const int64_t* ei, const int64_t* ej, const double* se_d,
                   const float* se_f, int64_t E64, int64_t m64, int64_t n64, double lam,
                   double gam, int iters, double damping, double eps, int64_t* assign,
                   double* hist, int* n_iter, double* belief_out) {
  if (!c) return fail(DE_ERR_ARG, "null ctx");
  if (E64 < 0 || m64 < 0 || n64 < 0 || E64 > (1ll << 30) || m64 > (1ll << 30) || n64 > (1ll << 30))
    return fail(DE_ERR_ARG, "edges %lld, m %lld, n %lld: each must be in [0, 2^30]",
                (long long)E64, (long long)m64, (long long)n64);
  if (iters < 0) return fail(DE_ERR_ARG, "iters %d < 0", iters);
  if (E64 > 0 && (!ei || !ej || (!se_d && !se_f))) return fail(DE_ERR_ARG, "null edge arrays");
  const int E = int(E64), m = int(m64), n = int(n64);
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  const size_t e8 = size_t(std::max(E, 1)) * 8;
  const size_t io = a256(e8) * 3 + a256(size_t(std::max(E, 1)) * 4);
  if (int rc = x.ed_io.ensure(io)) return rc;
  if (int rc = x.ed_scratch.ensure(edge_scratch_bytes(E, m, n, iters))) return rc;
  EdgeSet s;
  std::memset(&s, 0, sizeof(s));
  s.E = E;
  s.m = m;
  s.n = n;
  edge_carve(x.ed_scratch.ptr, E, m, n, iters, &s);
  Staging g{static_cast<uint8_t*>(x.ed_io.ptr)};
  s.ei64 = static_cast<const int64_t*>(g.put(ei, size_t(E) * 8, x.stream));
  s.ej64 = static_cast<const int64_t*>(g.put(ej, size_t(E) * 8, x.stream));
  double* se = static_cast<double*>(g.take(e8));
  if (se_d) {
    if (E) DE_HIP(hipMemcpyAsync(se, se_d, size_t(E) * 8, hipMemcpyHostToDevice, x.stream));
  } else if (E) {
    float* tmp = static_cast<float*>(g.put(se_f, size_t(E) * 4, x.stream));
    hipLaunchKernelGGL(k_f32_to_f64, dim3(gblocks(E)), dim3(kT), 0, x.stream, tmp, se, E);
  }
  DE_HIP(hipGetLastError());
  s.se = se;
  if (int rc = edge_solve(x, s, iters, lam, gam, damping, eps, true, true, x.stream)) return rc;
  int state[4];
  DE_HIP(hipMemcpy(state, s.state, 16, hipMemcpyDeviceToHost));
  if (n_iter) *n_iter = state[2];
  if (hist && state[2] > 0)
    DE_HIP(hipMemcpy(hist, s.hist, size_t(state[2]) * 8, hipMemcpyDeviceToHost));
  if (belief_out && E)
    DE_HIP(hipMemcpy(belief_out, s.bel, size_t(E) * 8, hipMemcpyDeviceToHost));
  if (assign && m) {
    std::vector<int> a(static_cast<size_t>(m));
    DE_HIP(hipMemcpy(a.data(), s.assign, size_t(m) * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < m; ++i) assign[i] = a[size_t(i)];
  }
  return DE_OK;
}

long long dense_edges(int W, int H, int hw, int hh, int dmin, int dmax, std::vector<long long>* st) {
  st->assign(1, 0);
  long long tot = 0;
  for (int d = dmin; d <= dmax; ++d) {
    const long long nx = std::max(0, W - 2 * hw - d);
    tot += nx * std::max(0, H - 2 * hh);
    st->push_back(tot);
  }
  return tot;
}

}  // namespace

extern "C" {

int de_masda_edges(de_ctx* c, const int64_t* ei, const int64_t* ej, const double* se,
                   int64_t n_edges, int64_t m, int64_t n, double lam, double gam, int iters,
                   double damping, double eps, int64_t* assign, double* hist, int* n_iter) {
  return run_edges_host(c, ei, ej, se, nullptr, n_edges, m, n, lam, gam, iters, damping, eps,
                        assign, hist, n_iter, nullptr);
}

int de_masda_edges_f32(de_ctx* c, const int64_t* ei, const int64_t* ej, const float* se,
                       int64_t n_edges, int64_t m, int64_t n, double lam, double gam, int iters,
                       double damping, double eps, int64_t* assign, double* hist, int* n_iter) {
  return run_edges_host(c, ei, ej, nullptr, se, n_edges, m, n, lam, gam, iters, damping, eps,
                        assign, hist, n_iter, nullptr);
}

int de_masda_matrix(de_ctx* c, const float* S, int64_t m64, int64_t n64, double lam, double gam,
                    int iters, double damping, double eps, int64_t* assign, double* hist,
                    int* n_iter) {
  if (!c || (!S && m64 * n64 > 0)) return fail(DE_ERR_ARG, "de_masda_matrix: null argument");
  if (m64 < 0 || n64 < 0 || m64 * n64 > (1ll << 31))
    return fail(DE_ERR_ARG, "matrix %lld x %lld", (long long)m64, (long long)n64);
  const int m = int(m64), n = int(n64);
  int E = 0;
  std::vector<int64_t> ei, ej;
  std::vector<double> se;
  {
    std::lock_guard<std::mutex> lk(c->c.mu);
    Ctx& x = c->c;
    DE_HIP(hipSetDevice(x.device));
    const size_t mn = size_t(m) * size_t(n);
    if (int rc = x.ed_io.ensure(a256(std::max<size_t>(mn, 1) * 4) + a256((size_t(m) + 1) * 4) +
                                a256(std::max<size_t>(mn, 1) * 8) * 3))
      return rc;
    uint8_t* p = static_cast<uint8_t*>(x.ed_io.ptr);
    float* dS = reinterpret_cast<float*>(p);
    p += a256(std::max<size_t>(mn, 1) * 4);
    int* cnt = reinterpret_cast<int*>(p);
    p += a256((size_t(m) + 1) * 4);
    int64_t* dei = reinterpret_cast<int64_t*>(p);
    p += a256(std::max<size_t>(mn, 1) * 8);
    int64_t* dej = reinterpret_cast<int64_t*>(p);
    p += a256(std::max<size_t>(mn, 1) * 8);
    double* dse = reinterpret_cast<double*>(p);
    if (m > 0 && n > 0) {
      DE_HIP(hipMemcpyAsync(dS, S, mn * 4, hipMemcpyHostToDevice, x.stream));
      hipLaunchKernelGGL(k_ed_zero_int, dim3(gblocks(m + 1)), dim3(kT), 0, x.stream, cnt, m + 1);
      hipLaunchKernelGGL(k_mat_count, dim3(m), dim3(kT), 0, x.stream, dS, m, n, cnt);
      hipLaunchKernelGGL(k_ed_scan, dim3(1), dim3(1024), 0, x.stream, cnt, m);
      hipLaunchKernelGGL(k_mat_edges, dim3(m), dim3(kT), 0, x.stream, dS, m, n, cnt, dei, dej, dse);
      DE_HIP(hipGetLastError());
      DE_HIP(hipMemcpyAsync(&E, cnt + m, 4, hipMemcpyDeviceToHost, x.stream));
      DE_HIP(hipStreamSynchronize(x.stream));
      ei.resize(size_t(E));
      ej.resize(size_t(E));
      se.resize(size_t(E));
      if (E) {
        DE_HIP(hipMemcpy(ei.data(), dei, size_t(E) * 8, hipMemcpyDeviceToHost));
        DE_HIP(hipMemcpy(ej.data(), dej, size_t(E) * 8, hipMemcpyDeviceToHost));
        DE_HIP(hipMemcpy(se.data(), dse, size_t(E) * 8, hipMemcpyDeviceToHost));
      }
    }
  }
  return run_edges_host(c, ei.data(), ej.data(), se.data(), nullptr, E, m, n, lam, gam, iters,
                        damping, eps, assign, hist, n_iter, nullptr);
}

int de_masda_decide(de_ctx* c, const int64_t* ei, const int64_t* ej, const double* se,
                    const double* belief, int64_t E64, int64_t m64, int64_t n64, double lam,
                    int64_t* assign) {
  if (!c) return fail(DE_ERR_ARG, "null ctx");
  if (E64 < 0 || m64 < 0 || n64 < 0 || E64 > (1ll << 30) || m64 > (1ll << 30) || n64 > (1ll << 30))
    return fail(DE_ERR_ARG, "decide: sizes out of range");
  if (E64 > 0 && (!ei || !ej || !se || !belief)) return fail(DE_ERR_ARG, "decide: null arrays");
  const int E = int(E64), m = int(m64), n = int(n64);
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  const size_t e8 = size_t(std::max(E, 1)) * 8;
  if (int rc = x.ed_io.ensure(a256(e8) * 3)) return rc;
  if (int rc = x.ed_scratch.ensure(edge_scratch_bytes(E, m, n, 1))) return rc;
  EdgeSet s;
  std::memset(&s, 0, sizeof(s));
  s.E = E;
  s.m = m;
  s.n = n;
  edge_carve(x.ed_scratch.ptr, E, m, n, 1, &s);
  Staging g{static_cast<uint8_t*>(x.ed_io.ptr)};
  s.ei64 = static_cast<const int64_t*>(g.put(ei, size_t(E) * 8, x.stream));
  s.ej64 = static_cast<const int64_t*>(g.put(ej, size_t(E) * 8, x.stream));
  s.se = static_cast<const double*>(g.put(se, size_t(E) * 8, x.stream));
  if (E) DE_HIP(hipMemcpyAsync(s.bel, belief, size_t(E) * 8, hipMemcpyHostToDevice, x.stream));
  if (int rc = edge_solve(x, s, 0, lam, lam, 0.0, 0.0, false, true, x.stream)) return rc;
  if (assign && m) {
    std::vector<int> a(static_cast<size_t>(m));
    DE_HIP(hipMemcpy(a.data(), s.assign, size_t(m) * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < m; ++i) assign[i] = a[size_t(i)];
  }
  return DE_OK;
}

int de_seg_max_excluding(de_ctx* c, const double* vals, const int64_t* idx, int64_t E64,
                         int64_t n64, double* out) {
  if (!c || (E64 > 0 && (!vals || !idx || !out))) return fail(DE_ERR_ARG, "null argument");
  if (E64 < 0 || n64 < 0 || E64 > (1ll << 30) || n64 > (1ll << 30))
    return fail(DE_ERR_ARG, "sizes out of range");
  const int E = int(E64), n = int(n64);
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  const size_t e8 = size_t(std::max(E, 1)) * 8;
  if (int rc = x.ed_io.ensure(a256(e8) * 3)) return rc;
  if (int rc = x.ed_scratch.ensure(edge_scratch_bytes(E, 1, n, 1))) return rc;
  EdgeSet s;
  std::memset(&s, 0, sizeof(s));
  s.E = E;
  s.m = 1;
  s.n = n;
  edge_carve(x.ed_scratch.ptr, E, 1, n, 1, &s);
  Staging g{static_cast<uint8_t*>(x.ed_io.ptr)};
  // segments as columns: ei = 0 for every element, ej = idx
  s.ej64 = static_cast<const int64_t*>(g.put(idx, size_t(E) * 8, x.stream));
  int64_t* zeros = static_cast<int64_t*>(g.take(e8));
  if (E) DE_HIP(hipMemsetAsync(zeros, 0, size_t(E) * 8, x.stream));
  s.ei64 = zeros;
  double* se = static_cast<double*>(g.take(e8));
  if (E) DE_HIP(hipMemsetAsync(se, 0, size_t(E) * 8, x.stream));
  s.se = se;
  if (int rc = edge_solve(x, s, 0, 0, 0, 0, 0, false, false, x.stream)) return rc;
  if (E == 0) return DE_OK;
  double* dv = s.rho;
  double* dout = s.beta;
  DE_HIP(hipMemcpyAsync(dv, vals, size_t(E) * 8, hipMemcpyHostToDevice, x.stream));
  hipLaunchKernelGGL(k_ed_segmax, dim3(gblocks(n)), dim3(kT), 0, x.stream, s, dv, dout);
  DE_HIP(hipGetLastError());
  DE_HIP(hipMemcpyAsync(out, dout, size_t(E) * 8, hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipStreamSynchronize(x.stream));
  return DE_OK;
}

int de_dense_problem(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H, int dmin,
                     int dmax, int hw, int hh, int64_t* ei, int64_t* ej, float* se,
                     int64_t cap, int64_t* n_edges) {
  if (!c || !L || !R || !n_edges) return fail(DE_ERR_ARG, "de_dense_problem: null argument");
  if (W < 1 || H < 1 || dmin < 0 || dmax < dmin)
    return fail(DE_ERR_ARG, "dense_problem: %dx%d, d in [%d, %d] (dmin >= 0)", W, H, dmin, dmax);
  if (!((hw == 3 && hh == 3) || (hw == 4 && hh == 3) || (hw == 2 && hh == 2) || (hw == 1 && hh == 1)))
    return fail(DE_ERR_ARG, "census window %dx%d", 2 * hw + 1, 2 * hh + 1);
  std::vector<long long> st;
  const long long E = dense_edges(W, H, hw, hh, dmin, dmax, &st);
  *n_edges = E;
  if (!ei && !ej && !se) return DE_OK;   // size query
  if (E > cap) return fail(DE_ERR_ARG, "dense_problem: %lld edges exceed capacity %lld", E,
                           (long long)cap);
  if (E > (1ll << 31)) return fail(DE_ERR_UNSUPPORTED, "dense_problem: %lld edges", E);
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  const size_t WH = size_t(W) * H;
  const int nd = dmax - dmin + 1;
  const size_t io = a256(WH) * 2 + a256(WH * 8) * 2 + a256(size_t(nd + 1) * 8) +
                    a256(size_t(std::max(E, 1ll)) * 8) * 2 + a256(size_t(std::max(E, 1ll)) * 4);
  if (int rc = x.ed_io.ensure(io)) return rc;
  Staging g{static_cast<uint8_t*>(x.ed_io.ptr)};
  const uint8_t* dL = static_cast<const uint8_t*>(g.put(L, WH, x.stream));
  const uint8_t* dR = static_cast<const uint8_t*>(g.put(R, WH, x.stream));
  DpDesc p;
  p.cl = static_cast<uint64_t*>(g.take(WH * 8));
  p.cr = static_cast<uint64_t*>(g.take(WH * 8));
  p.start = static_cast<const long long*>(g.put(st.data(), size_t(nd + 1) * 8, x.stream));
  p.ei = static_cast<int64_t*>(g.take(size_t(std::max(E, 1ll)) * 8));
  p.ej = static_cast<int64_t*>(g.take(size_t(std::max(E, 1ll)) * 8));
  p.se = static_cast<float*>(g.take(size_t(std::max(E, 1ll)) * 4));
  p.W = W;
  p.H = H;
  p.hw = hw;
  p.hh = hh;
  p.dmin = dmin;
  p.nd = nd;
  DE_HIP(launch_census_plane(dL, W, H, hw, hh, const_cast<uint64_t*>(p.cl), x.stream));
  DE_HIP(launch_census_plane(dR, W, H, hw, hh, const_cast<uint64_t*>(p.cr), x.stream));
  if (E > 0) {
    hipLaunchKernelGGL(k_dense_problem, dim3(gblocks(E)), dim3(kT), 0, x.stream, p, E);
    DE_HIP(hipGetLastError());
    if (ei) DE_HIP(hipMemcpyAsync(ei, p.ei, size_t(E) * 8, hipMemcpyDeviceToHost, x.stream));
    if (ej) DE_HIP(hipMemcpyAsync(ej, p.ej, size_t(E) * 8, hipMemcpyDeviceToHost, x.stream));
    if (se) DE_HIP(hipMemcpyAsync(se, p.se, size_t(E) * 4, hipMemcpyDeviceToHost, x.stream));
  }
  DE_HIP(hipStreamSynchronize(x.stream));
  return DE_OK;
}

int de_dense_masda(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H, int dmin,
                   int dmax, int hw, int hh, double lam, double gam, int iters, double damping,
                   double eps, float* disp, double* hist, int* n_iter) {
  if (!c || !L || !R || !disp) return fail(DE_ERR_ARG, "de_dense_masda: null argument");
  int64_t E = 0;
  if (int rc = de_dense_problem(c, L, R, W, H, dmin, dmax, hw, hh, nullptr, nullptr, nullptr, 0, &E))
    return rc;
  if (E > (1ll << 30)) return fail(DE_ERR_UNSUPPORTED, "dense_masda: %lld edges", (long long)E);
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  const size_t WH = size_t(W) * H;
  const int nd = dmax - dmin + 1;
  std::vector<long long> st;
  dense_edges(W, H, hw, hh, dmin, dmax, &st);
  const size_t Ez = size_t(std::max<int64_t>(E, 1));
  const size_t io = a256(WH) * 2 + a256(WH * 8) * 2 + a256(size_t(nd + 1) * 8) + a256(Ez * 8) * 3 +
                    a256(Ez * 4) + a256(WH * 4);
  if (int rc = x.ed_io.ensure(io)) return rc;
  const int m = int(WH);
  if (int rc = x.ed_scratch.ensure(edge_scratch_bytes(int(E), m, m, iters))) return rc;
  Staging g{static_cast<uint8_t*>(x.ed_io.ptr)};
  const uint8_t* dL = static_cast<const uint8_t*>(g.put(L, WH, x.stream));
  const uint8_t* dR = static_cast<const uint8_t*>(g.put(R, WH, x.stream));
  DpDesc p;
  p.cl = static_cast<uint64_t*>(g.take(WH * 8));
  p.cr = static_cast<uint64_t*>(g.take(WH * 8));
  p.start = static_cast<const long long*>(g.put(st.data(), size_t(nd + 1) * 8, x.stream));
  p.ei = static_cast<int64_t*>(g.take(Ez * 8));
  p.ej = static_cast<int64_t*>(g.take(Ez * 8));
  double* se = static_cast<double*>(g.take(Ez * 8));
  p.se = static_cast<float*>(g.take(Ez * 4));
  float* ddisp = static_cast<float*>(g.take(WH * 4));
  p.W = W;
  p.H = H;
  p.hw = hw;
  p.hh = hh;
  p.dmin = dmin;
  p.nd = nd;
  DE_HIP(launch_census_plane(dL, W, H, hw, hh, const_cast<uint64_t*>(p.cl), x.stream));
  DE_HIP(launch_census_plane(dR, W, H, hw, hh, const_cast<uint64_t*>(p.cr), x.stream));
  if (E > 0) {
    hipLaunchKernelGGL(k_dense_problem, dim3(gblocks(E)), dim3(kT), 0, x.stream, p, (long long)E);
    hipLaunchKernelGGL(k_f32_to_f64, dim3(gblocks(E)), dim3(kT), 0, x.stream, p.se, se, int(E));
    DE_HIP(hipGetLastError());
  }
  EdgeSet s;
  std::memset(&s, 0, sizeof(s));
  s.E = int(E);
  s.m = m;
  s.n = m;
  edge_carve(x.ed_scratch.ptr, int(E), m, m, iters, &s);
  s.ei64 = p.ei;
  s.ej64 = p.ej;
  s.se = se;
  if (int rc = edge_solve(x, s, iters, lam, gam, damping, eps, true, true, x.stream)) return rc;
  hipLaunchKernelGGL(k_assign_disp, dim3(gblocks(m)), dim3(kT), 0, x.stream, s.assign, W, m, ddisp);
  DE_HIP(hipGetLastError());
  int state[4];
  DE_HIP(hipMemcpyAsync(state, s.state, 16, hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipMemcpyAsync(disp, ddisp, WH * 4, hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipStreamSynchronize(x.stream));
  if (n_iter) *n_iter = state[2];
  if (hist && state[2] > 0)
    DE_HIP(hipMemcpy(hist, s.hist, size_t(state[2]) * 8, hipMemcpyDeviceToHost));
  return DE_OK;
}

int de_census_points(de_ctx* c, const uint8_t* img, int W, int H, const int64_t* xy,
                     int64_t n, int hw, int hh, uint64_t* out) {
  if (!c || !img || (n > 0 && (!xy || !out))) return fail(DE_ERR_ARG, "de_census_points: null argument");
  if (W < 1 || H < 1 || n < 0 || n > (1ll << 31) || hw < 0 || hh < 0)
    return fail(DE_ERR_ARG, "census: image %dx%d, %lld points, window %dx%d", W, H, (long long)n,
                2 * hw + 1, 2 * hh + 1);
  if ((2 * hw + 1) * (2 * hh + 1) - 1 > 64)
    return fail(DE_ERR_UNSUPPORTED, "census: a %dx%d window has more than 64 bits", 2 * hw + 1,
                2 * hh + 1);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t x = xy[2 * i], y = xy[2 * i + 1];
    if (x - hw < -int64_t(W) || x + hw >= W || y - hh < -int64_t(H) || y + hh >= H)
      return fail(DE_ERR_ARG, "census: point %lld (%lld, %lld) indexes outside the %dx%d image",
                  (long long)i, (long long)x, (long long)y, W, H);
  }
  if (n == 0) return DE_OK;
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  const size_t WH = size_t(W) * H;
  if (int rc = x.ed_io.ensure(a256(WH) + a256(size_t(n) * 16) + a256(size_t(n) * 8))) return rc;
  Staging g{static_cast<uint8_t*>(x.ed_io.ptr)};
  const uint8_t* dimg = static_cast<const uint8_t*>(g.put(img, WH, x.stream));
  const int64_t* dxy = static_cast<const int64_t*>(g.put(xy, size_t(n) * 16, x.stream));
  uint64_t* dout = static_cast<uint64_t*>(g.take(size_t(n) * 8));
  hipLaunchKernelGGL(k_census_points, dim3(gblocks(n)), dim3(kT), 0, x.stream, dimg, W, H, dxy, n,
                     hw, hh, dout);
  DE_HIP(hipGetLastError());
  DE_HIP(hipMemcpyAsync(out, dout, size_t(n) * 8, hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipStreamSynchronize(x.stream));
  return DE_OK;
}

int de_hamming(de_ctx* c, const uint64_t* a, const uint64_t* b, int64_t n, int32_t* out) {
  if (!c || (n > 0 && (!a || !b || !out))) return fail(DE_ERR_ARG, "de_hamming: null argument");
  if (n < 0 || n > (1ll << 32)) return fail(DE_ERR_ARG, "hamming: %lld elements", (long long)n);
  if (n == 0) return DE_OK;
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  if (int rc = x.ed_io.ensure(a256(size_t(n) * 8) * 2 + a256(size_t(n) * 4))) return rc;
  Staging g{static_cast<uint8_t*>(x.ed_io.ptr)};
  const uint64_t* da = static_cast<const uint64_t*>(g.put(a, size_t(n) * 8, x.stream));
  const uint64_t* db = static_cast<const uint64_t*>(g.put(b, size_t(n) * 8, x.stream));
  int32_t* dout = static_cast<int32_t*>(g.take(size_t(n) * 4));
  hipLaunchKernelGGL(k_hamming, dim3(gblocks(n)), dim3(kT), 0, x.stream, da, db, n, dout);
  DE_HIP(hipGetLastError());
  DE_HIP(hipMemcpyAsync(out, dout, size_t(n) * 4, hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipStreamSynchronize(x.stream));
  return DE_OK;
}

int de_build_problem(de_ctx* c, const void* pl, const uint64_t* dl, int64_t m, const void* pr,
                     const uint64_t* dr, int64_t n, int fp64, double max_dy, double dmin,
                     double dmax, double sigma_y, double w_y, float* S) {
  if (!c || (m > 0 && (!pl || !dl)) || (n > 0 && (!pr || !dr)) || (m > 0 && n > 0 && !S))
    return fail(DE_ERR_ARG, "de_build_problem: null argument");
  if (m < 0 || n < 0 || m > (1 << 30) || n > (1 << 30) || m * n > (1ll << 32))
    return fail(DE_ERR_ARG, "build_problem: %lld x %lld", (long long)m, (long long)n);
  if (m == 0 || n == 0) return DE_OK;
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  const size_t ts = fp64 ? 8 : 4;
  const size_t mn = size_t(m) * size_t(n);
  if (int rc = x.ed_io.ensure(a256(size_t(m) * 2 * ts) + a256(size_t(n) * 2 * ts) +
                              a256(size_t(m) * 8) + a256(size_t(n) * 8) + a256(mn * 4)))
    return rc;
  Staging g{static_cast<uint8_t*>(x.ed_io.ptr)};
  const void* dpl = g.put(pl, size_t(m) * 2 * ts, x.stream);
  const void* dpr = g.put(pr, size_t(n) * 2 * ts, x.stream);
  const uint64_t* ddl = static_cast<const uint64_t*>(g.put(dl, size_t(m) * 8, x.stream));
  const uint64_t* ddr = static_cast<const uint64_t*>(g.put(dr, size_t(n) * 8, x.stream));
  float* dS = static_cast<float*>(g.take(mn * 4));
  if (fp64)
    hipLaunchKernelGGL(k_build_problem<double>, dim3(gblocks((long long)mn)), dim3(kT), 0,
                       x.stream, static_cast<const double*>(dpl), ddl, int(m),
                       static_cast<const double*>(dpr), ddr, int(n), max_dy, dmin, dmax, sigma_y,
                       w_y, dS);
  else
    hipLaunchKernelGGL(k_build_problem<float>, dim3(gblocks((long long)mn)), dim3(kT), 0,
                       x.stream, static_cast<const float*>(dpl), ddl, int(m),
                       static_cast<const float*>(dpr), ddr, int(n), float(max_dy), float(dmin),
                       float(dmax), float(sigma_y), float(w_y), dS);
  DE_HIP(hipGetLastError());
  DE_HIP(hipMemcpyAsync(S, dS, mn * 4, hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipStreamSynchronize(x.stream));
  return DE_OK;
}

}  // extern "C"
