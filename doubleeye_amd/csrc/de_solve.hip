// Dense MASDA row solver on gfx950: one workgroup per image row.
//
// Restates core/include/doubleeye/dense_solve.hpp:185-452 (solve_row_sparse,
// blockwise K = 2 path) plus the candidate planes of core/tools/de_dense.cpp:
// 1539-1560 and the DEDD quantisation/confidence of core/tools/de_dense_cuda.cu:
// 731-745 and dense_solve.hpp:158-171.
//
// Compiled with -ffp-contract=off: every float expression is evaluated in the
// reference's order with one rounding per operation, so the messages, beliefs,
// margins and sub-pixel offsets are bit-identical to the x86 reference build.
//
// Parallel forms of the two sequential pieces:
//  * claimant buckets: an LDS counting sort whose within-bucket order is
//    arbitrary. The beta update only needs max-excluding-self over a bucket,
//    which does not depend on order (a unique maximum is excluded from itself;
//    tied maxima make the runner-up equal to the maximum). The reverse match
//    takes the first maximum in ascending x explicitly.
//  * the greedy claim in std::sort order: each pixel has exactly one target
//    right pixel, so "first in descending-belief order claims it" is a per-target
//    argmax of belief. Exactly equal beliefs competing for one target are
//    ordered by std::sort's unstable partition in the reference; here the
//    smaller x wins. Such pixels are the documented exact-tie pixels.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "de_batch.h"
#include "de_common.h"

namespace de {

namespace {

constexpr int kSolveThreads = 512;   // 8 waves: two and more rows resident per CU

__device__ __forceinline__ float fmax_ref(float a, float b) {   // std::max(a, b)
  return (a < b) ? b : a;
}

__device__ __forceinline__ uint32_t ordered(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

struct Top2 {
  float b1 = kNone, b2 = kNone;
  int i1 = -1;
  __device__ void push(float v, int i) {
    if (v > b1) { b2 = b1; b1 = v; i1 = i; }
    else if (v > b2) { b2 = v; }
  }
  __device__ float excl(int i) const { return i == i1 ? b2 : b1; }
};

constexpr int kBucketRegs = 4;                      // bucket items kept in registers

// LDS layout for one row of width W (bytes): see solve_lds_bytes().
struct SolveSmem {
  float* cs;        // [2W]
  int16_t* cd;      // [2W]
  uint8_t* cn;      // [W]
  float* beta;      // [2W]   (reused as claim keys after the iterations)
  float* rho;       // [2W]
  int* bstart;      // [W + 1]
  int* cursor;      // [W]
  int16_t* items;   // [2W]
  int16_t* rev;     // [W]
  float* best;      // [W]
  int8_t* bestj;    // [W]
  int* scan;        // [kSolveThreads]
};

__device__ SolveSmem carve(uint8_t* base, int W) {
  SolveSmem s;
  size_t o = 0;
  auto take = [&](size_t bytes) { uint8_t* p = base + o; o += (bytes + 15) & ~size_t(15); return p; };
  s.cs = reinterpret_cast<float*>(take(8 * size_t(W)));
  s.beta = reinterpret_cast<float*>(take(8 * size_t(W)));
  s.rho = reinterpret_cast<float*>(take(8 * size_t(W)));
  s.bstart = reinterpret_cast<int*>(take(4 * size_t(W + 1)));
  s.cursor = reinterpret_cast<int*>(take(4 * size_t(W)));
  s.best = reinterpret_cast<float*>(take(4 * size_t(W)));
  s.scan = reinterpret_cast<int*>(take(4 * size_t(kSolveThreads)));
  s.cd = reinterpret_cast<int16_t*>(take(4 * size_t(W)));
  s.items = reinterpret_cast<int16_t*>(take(4 * size_t(W)));
  s.rev = reinterpret_cast<int16_t*>(take(2 * size_t(W)));
  s.cn = take(size_t(W));
  s.bestj = reinterpret_cast<int8_t*>(take(size_t(W)));
  return s;
}

__device__ __forceinline__ float key_score(uint32_t key) {
  return float(key_v(key)) * (1.f / float(kOne));
}

__device__ __forceinline__ float nb_score(uint32_t h) {
  return (h == 0xFFFFu) ? kNone : float(int(h) - kOff) * (1.f / float(kOne));
}

// confidence(), dense_solve.hpp:158-171, with the device's logf/expf.
__device__ float confidence_dev(float s1, float s2, float lambda, float lrc, float lrc_tol) {
  if (!(s1 > -1e29f)) return 0.f;
  const float alt = (s2 > -1e29f) ? fmax_ref(s2, lambda) : lambda;
  const float x = logf(fmax_ref((1.f - alt) / fmax_ref(1.f - s1, 1e-6f), 1e-6f));
  const float logit = 0.692684f + 7.936751f * x;
  const float cl = fmax_ref(-30.f, fminf(30.f, logit));
  const float p = 1.f / (1.f + expf(-cl));
  return (lrc > lrc_tol) ? fminf(p, 0.35f) : p;
}

}  // namespace

// KP: pixels (and buckets) per thread, at least ceil(W / kSolveThreads) for
// every frame of the launch.
template <int KP>
__global__ __launch_bounds__(kSolveThreads) void k_solve(BatchView b) {
  constexpr int kPerThread = KP;
  extern __shared__ __align__(16) uint8_t smem_raw[];
  const int item = blockIdx.x;
  const int f = find_frame(b, kStageSolve, item);
  const FrameDesc F = b.frames[f];
  const int y = item - b.start[kStageSolve * (b.n + 1) + f];
  const int W = F.W, dmin = F.dmin;
  const int tid = threadIdx.x, nt = blockDim.x;
  const float lambda = F.lambda, gamma = F.gamma, damping = F.damping;
  const float keep = 1.f - damping;
  SolveSmem S = carve(smem_raw, W);
  const DE_CONST uint4* recs = cptr(reinterpret_cast<const uint4*>(F.cand)) + size_t(y) * W;
  DE_GLOBAL float* g_disp = gptr(F.disp);
  DE_GLOBAL float* g_margin = gptr(F.margin);
  const size_t row = size_t(y) * W;

  // --- candidates, margin, bucket counts ---
  for (int x = tid; x < W + 1; x += nt) S.bstart[x] = 0;
  __syncthreads();
  for (int x = tid; x < W; x += nt) {
    const uint4 r = recs[x];
    int n = 0;
    S.cs[2 * x] = S.cs[2 * x + 1] = kNone;
    S.cd[2 * x] = S.cd[2 * x + 1] = -1;
    if (r.x) {
      S.cs[2 * x] = key_score(r.x); S.cd[2 * x] = int16_t(key_k(r.x)); n = 1;
      if (r.y) { S.cs[2 * x + 1] = key_score(r.y); S.cd[2 * x + 1] = int16_t(key_k(r.y)); n = 2; }
    }
    S.cn[x] = uint8_t(n);
    const float alt = (n < 2) ? lambda : fmax_ref(S.cs[2 * x + 1], lambda);
    g_margin[row + x] = (n == 0) ? 0.f : S.cs[2 * x] - alt;
    S.beta[2 * x] = S.beta[2 * x + 1] = 0.f;
    S.rho[2 * x] = S.rho[2 * x + 1] = 0.f;
    for (int j = 0; j < n; ++j) {
      const int xr = x - (dmin + S.cd[2 * x + j]);
      if (xr >= 0 && xr < W) atomicAdd(&S.bstart[xr], 1);
    }
  }
  __syncthreads();

  // --- exclusive scan of the counts: chunk per thread, inclusive scan within
  // each wave through lane shuffles, then over the wave totals ---
  {
    const int per = (W + nt - 1) / nt;
    const int lo = min(W, tid * per), hi = min(W, lo + per);
    const int lane = tid & 63, wave = tid >> 6, nwave = nt >> 6;
    int sum = 0;
    for (int i = lo; i < hi; ++i) sum += S.bstart[i];
    int v = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int u = __shfl_up(v, off);
      if (lane >= off) v += u;
    }
    if (lane == 63) S.scan[wave] = v;
    __syncthreads();
    if (wave == 0) {
      int w = lane < nwave ? S.scan[lane] : 0;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(w, off);
        if (lane >= off) w += u;
      }
      if (lane < nwave) S.scan[lane] = w;
    }
    __syncthreads();
    int run = (wave > 0 ? S.scan[wave - 1] : 0) + v - sum;
    for (int i = lo; i < hi; ++i) { const int c = S.bstart[i]; S.bstart[i] = run; run += c; }
    if (tid == nt - 1) S.bstart[W] = S.scan[nwave - 1];
  }
  __syncthreads();
  for (int x = tid; x < W; x += nt) S.cursor[x] = S.bstart[x];
  __syncthreads();
  for (int x = tid; x < W; x += nt) {
    const int n = S.cn[x];
    for (int j = 0; j < n; ++j) {
      const int xr = x - (dmin + S.cd[2 * x + j]);
      if (xr >= 0 && xr < W) S.items[atomicAdd(&S.cursor[xr], 1)] = int16_t(2 * x + j);
    }
  }
  __syncthreads();

  // --- reverse match: best claimant by score, first (smallest x) of equals ---
  for (int xr = tid; xr < W; xr += nt) {
    float bv = kNone;
    int bx = 0x7FFFFFFF, bd = -1;
    for (int p = S.bstart[xr]; p < S.bstart[xr + 1]; ++p) {
      const int idx = S.items[p];
      const float v = S.cs[idx];
      const int xx = idx >> 1;
      if (v > bv || (v == bv && xx < bx)) { bv = v; bx = xx; bd = S.cd[idx]; }
    }
    S.rev[xr] = int16_t(bd);
  }

  // --- message passing ---
  // Thread tid owns pixels x = tid + k*nt and buckets xr = tid + k*nt. A
  // pixel's candidates, beta and rho live in its owner's registers across
  // the iterations; rho goes to LDS for the bucket owners, which publish each
  // bucket's top-2 of rho (best, runner-up, item index of the first best);
  // the pixel owner then updates its beta from that. Item indices are unique
  // in a bucket, so excluding by index equals excluding by position, and the
  // float expressions are the reference's, operation for operation.
  //
  // The three phases run branch-free: absent candidates, pixels at or beyond
  // W and unused bucket slots compute on neutral operands (slot 0, kNone) and
  // their results are dropped by selects, so every lane issues the same
  // instructions. A push of kNone leaves a top-2 as it is.
  float2* const bt = reinterpret_cast<float2*>(S.beta);   // [W] bucket best, runner-up rho
  int* const bti = S.cursor;                              // [W] item index of the best (-1: empty)
  float c_s[kPerThread][2], c_b[kPerThread][2], c_r[kPerThread][2];
  int c_xr[kPerThread][2], c_n[kPerThread];
  int b_lo[kPerThread], b_cnt[kPerThread], b_it[kPerThread][kBucketRegs];
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int x = tid + k * nt;
    c_n[k] = 0;
    b_cnt[k] = 0;
    b_lo[k] = 0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      c_s[k][j] = kNone; c_b[k][j] = 0.f; c_r[k][j] = 0.f; c_xr[k][j] = -1;
    }
#pragma unroll
    for (int i = 0; i < kBucketRegs; ++i) b_it[k][i] = 0;
    if (x < W) {
      const int n = S.cn[x];
      c_n[k] = n;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        if (j < n) {
          c_s[k][j] = S.cs[2 * x + j];
          const int xr = x - (dmin + S.cd[2 * x + j]);
          c_xr[k][j] = (xr >= 0 && xr < W) ? xr : -1;
        }
      }
      const int lo = S.bstart[x], cnt = S.bstart[x + 1] - lo;
      b_lo[k] = lo;
      b_cnt[k] = cnt;
#pragma unroll
      for (int i = 0; i < kBucketRegs; ++i) if (i < cnt) b_it[k][i] = S.items[lo + i];
    }
  }
  for (int it = 0; it < F.iters; ++it) {
    // rho of every pixel from its own betas. Top2::push of beta 0 then beta 1
    // gives excl(0) = beta 1 (if present and above kNone, else kNone) and
    // excl(1) = beta 0 (if above kNone, else kNone).
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
      const int n = c_n[k];
      const int x = tid + k * nt;
      const float e0 = (n > 1 && c_b[k][1] > kNone) ? c_b[k][1] : kNone;
      const float e1 = (c_b[k][0] > kNone) ? c_b[k][0] : kNone;
      const float t0 = c_s[k][0] - fmax_ref(lambda, e0);
      const float t1 = c_s[k][1] - fmax_ref(lambda, e1);
      const float r0 = keep * t0 + damping * c_r[k][0];
      const float r1 = keep * t1 + damping * c_r[k][1];
      c_r[k][0] = n > 0 ? r0 : c_r[k][0];
      c_r[k][1] = n > 1 ? r1 : c_r[k][1];
      if (x < W) reinterpret_cast<float2*>(S.rho)[x] = make_float2(c_r[k][0], c_r[k][1]);
    }
    __syncthreads();
    // each bucket's top-2 of rho (slots past the count push kNone)
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
      const int xr = tid + k * nt;
      const int cnt = b_cnt[k];
      float t1 = kNone, t2 = kNone;
      int ti = -1;
      auto push = [&](float v, int i) {
        const bool g1 = v > t1, g2 = v > t2;
        t2 = g1 ? t1 : (g2 ? v : t2);
        t1 = g1 ? v : t1;
        ti = g1 ? i : ti;
      };
#pragma unroll
      for (int i = 0; i < kBucketRegs; ++i) {
        const float rv = S.rho[b_it[k][i]];
        push(i < cnt ? rv : kNone, b_it[k][i]);
      }
      for (int i = kBucketRegs; i < cnt; ++i) {
        const int idx = S.items[b_lo[k] + i];
        push(S.rho[idx], idx);
      }
      if (xr < W) {
        bt[xr] = make_float2(t1, t2);
        bti[xr] = ti;
      }
    }
    __syncthreads();
    // beta of every bucketed candidate from its bucket's top-2
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
      const int x = tid + k * nt;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int xr = c_xr[k][j];          // -1: absent or outside the row
        const int xq = max(xr, 0);
        const float2 tb = bt[xq];
        const float ex = bti[xq] == 2 * x + j ? tb.y : tb.x;
        const float nb = keep * (c_s[k][j] - fmax_ref(gamma, ex)) + damping * c_b[k][j];
        c_b[k][j] = xr >= 0 ? nb : c_b[k][j];
      }
    }
  }
  // final betas and rhos back to LDS for the belief pass (the bucket top-2
  // arrays share the beta storage: every thread is past its last read)
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int x = tid + k * nt;
    if (x >= W) continue;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      S.beta[2 * x + j] = c_b[k][j];
      S.rho[2 * x + j] = c_r[k][j];
    }
  }
  __syncthreads();

  // --- belief, gates, claim ---
  for (int x = tid; x < W; x += nt) {
    float bst = kNone;
    int bj = -1;
    for (int j = 0; j < S.cn[x]; ++j) {
      const float c = S.cs[2 * x + j];
      if (c <= lambda) continue;
      const float bel = (F.iters > 0) ? (S.beta[2 * x + j] + S.rho[2 * x + j] - c) : c;
      if (bel > bst) { bst = bel; bj = j; }
    }
    if (bj >= 0 && F.min_margin > 0.f && g_margin[row + x] < F.min_margin) bj = -1;
    if (bj >= 0 && F.min_ratio > 0.f) {
      const float s2 = (S.cn[x] > 1) ? fmax_ref(S.cs[2 * x + 1], lambda) : lambda;
      const float c1 = fmax_ref(1.f - S.cs[2 * x], 1e-6f);
      if ((1.f - s2) / c1 < F.min_ratio) bj = -1;
    }
    if (bj >= 0 && x - (dmin + S.cd[2 * x + bj]) < 0) bj = -1;
    S.best[x] = bst;
    S.bestj[x] = int8_t(bj);
  }
  __syncthreads();
  unsigned long long* claim = reinterpret_cast<unsigned long long*>(S.beta);
  for (int x = tid; x < W; x += nt) claim[x] = 0ull;
  __syncthreads();
  for (int x = tid; x < W; x += nt) {
    const int bj = S.bestj[x];
    if (bj < 0) continue;
    const int xr = x - (dmin + S.cd[2 * x + bj]);
    const unsigned long long key =
        (static_cast<unsigned long long>(ordered(S.best[x])) << 32) | (0xFFFFFFFFu - uint32_t(x));
    atomicMax(&claim[xr], key);
  }
  __syncthreads();

  // --- outputs ---
  const float q = 1.f / float(kOne);
  const size_t WH = size_t(W) * F.H;
  for (int x = tid; x < W; x += nt) {
    const int bj = S.bestj[x];
    float out = __int_as_float(0x7fc00000);
    if (bj >= 0) {
      const int k = S.cd[2 * x + bj];
      const int d = dmin + k;
      const int xr = x - d;
      const uint32_t mine = 0xFFFFFFFFu - uint32_t(x);
      if (uint32_t(claim[xr] & 0xFFFFFFFFull) == mine) {
        float off = 0.f;
        if (F.subpixel && bj == 0) {
          const uint32_t nb = recs[x].z;
          const float cm = nb_score(nb & 0xFFFFu), cp = nb_score(nb >> 16);
          if (cm > -1e29f && cp > -1e29f) {
            const float c0 = S.cs[2 * x];
            if (F.fit_eq) {
              const float lower = cm < cp ? cm : cp;
              const float den = c0 - lower;
              if (den > 1e-6f) off = 0.5f * (cp - cm) / den;
            } else {
              const float den = 2.f * c0 - cm - cp;
              if (den > 1e-6f) off = 0.5f * (cp - cm) / den;
            }
            if (off > 0.5f) off = 0.5f;
            if (off < -0.5f) off = -0.5f;
          }
        }
        out = float(d) + off;
      }
    }
    g_disp[row + x] = out;
    const bool fin = out == out;
    float lrc2 = __int_as_float(0x7fc00000);
    if (fin) {
      const int xr = x - int(roundf(out));
      if (xr >= 0 && xr < W && S.rev[xr] >= 0) lrc2 = fabsf(out - float(dmin + S.rev[xr]));
    }
    const int n = S.cn[x];
    if (F.cand6) {
      DE_GLOBAL float* c6 = gptr(F.cand6) + row + x;
      c6[0] = n > 0 ? S.cs[2 * x] : kNone;
      c6[WH] = n > 1 ? S.cs[2 * x + 1] : kNone;
      c6[2 * WH] = n > 0 ? float(dmin + S.cd[2 * x]) : kNone;
      c6[3 * WH] = n > 1 ? float(dmin + S.cd[2 * x + 1]) : kNone;
      c6[4 * WH] = kNone;   // exact reverse match (--lrc) is written by k_lrc_exact
      c6[5 * WH] = lrc2;
    }
    if (F.q4) {
      gptr(F.q4)[row + x] = (fin && out > 0.f) ? int16_t(roundf(out * 16.0f)) : int16_t(-32768);
    }
    if (F.conf) {
      const bool has = fin && out > 0.f;
      gptr(F.conf)[row + x] = has ? uint8_t(roundf(confidence_dev(n > 0 ? S.cs[2 * x] : kNone,
                                                            n > 1 ? S.cs[2 * x + 1] : kNone,
                                                            lambda, (lrc2 == lrc2) ? lrc2 : 0.f,
                                                            1.f) * 255.f))
                            : uint8_t(0);
    }
    (void)q;
  }
}

size_t solve_lds_bytes(int W) {
  auto r = [](size_t v) { return (v + 15) & ~size_t(15); };
  return r(8 * size_t(W)) * 3 + r(4 * size_t(W + 1)) + r(4 * size_t(W)) * 2 +
         r(4 * size_t(kSolveThreads)) + r(4 * size_t(W)) * 2 + r(2 * size_t(W)) +
         r(size_t(W)) * 2;
}

// maxW: widest row of the launch (at most kMaxW).
hipError_t launch_solve(const BatchView& b, int items, size_t lds, int maxW, hipStream_t s) {
  const int kp = (maxW + kSolveThreads - 1) / kSolveThreads;
  if (kp <= 1) hipLaunchKernelGGL(k_solve<1>, dim3(items), dim3(kSolveThreads), lds, s, b);
  else if (kp == 2) hipLaunchKernelGGL(k_solve<2>, dim3(items), dim3(kSolveThreads), lds, s, b);
  else if (kp <= 4) hipLaunchKernelGGL(k_solve<4>, dim3(items), dim3(kSolveThreads), lds, s, b);
  else if (kp <= 8) hipLaunchKernelGGL(k_solve<8>, dim3(items), dim3(kSolveThreads), lds, s, b);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace de
