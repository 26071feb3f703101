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

constexpr int kSolveThreads = 512;   // 8 waves (16 for rows of 513..1024 pixels)

__device__ __forceinline__ float fmax_ref(float a, float b) {   // std::max(a, b)
  return (a < b) ? b : a;
}

__device__ __forceinline__ uint32_t ordered(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

constexpr int kBucketRegs = 4;                      // bucket items kept in registers

// LDS layout for one row of width W (bytes): see solve_lds_bytes().
struct SolveSmem {
  float* rho;       // [2W]  rho per candidate (2x + j); claim keys [W] (u64) after the iterations
  float2* bt;       // [W]   bucket best and runner-up rho
  unsigned long long* revk;   // [W] reverse-match key per right pixel (0: no claimant)
  int* bstart;      // [W + 1]
  int* cursor;      // [W]   placement cursors; item index of each bucket's best during the iterations
  int16_t* items;   // [2W]
  int* scan;        // [kSolveThreads] (wave totals: at most 16 used)
};

__device__ SolveSmem carve(uint8_t* base, int W) {
  SolveSmem s;
  size_t o = 0;
  auto take = [&](size_t bytes) { uint8_t* p = base + o; o += (bytes + 15) & ~size_t(15); return p; };
  s.rho = reinterpret_cast<float*>(take(8 * size_t(W)));
  s.bt = reinterpret_cast<float2*>(take(8 * size_t(W)));
  s.revk = reinterpret_cast<unsigned long long*>(take(8 * size_t(W)));
  s.bstart = reinterpret_cast<int*>(take(4 * size_t(W + 1)));
  s.cursor = reinterpret_cast<int*>(take(4 * size_t(W)));
  s.items = reinterpret_cast<int16_t*>(take(4 * size_t(W)));
  s.scan = reinterpret_cast<int*>(take(4 * size_t(kSolveThreads)));
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

// KP: pixels (and buckets) per thread and NT threads, KP * NT at least the
// widest row of the launch. Thread tid owns pixels x = tid + k*nt and the
// buckets (right pixels) of the same indices, k < KP; everything about its
// own pixels stays in its registers from the candidate parse to the outputs.
template <int KP, int NT = kSolveThreads>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(KP <= 4 ? 4 : 2)))
void k_solve(BatchView b) {
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

  for (int x = tid; x < W + 1; x += nt) S.bstart[x] = 0;
  for (int x = tid; x < W; x += nt) S.revk[x] = 0ull;
  __syncthreads();

  // --- candidates and margin into registers; bucket counts; reverse match ---
  // A candidate j of pixel x claims right pixel xr = x - (dmin + k_j) when it
  // lies in the row. The reverse match of xr is its claimant of highest score,
  // the smallest x among equal scores (scores are never -0, so the ordered
  // bits compare as the floats do): one 64-bit atomic max per claim over
  // (score, 0xFFFF - x, k). x is unique within a bucket (a pixel's two
  // candidates have different k).
  // The margin, (n == 0) ? 0 : s1 - max(s2, lambda) with lambda for a missing s2.
  auto margin_of = [&](int n, float s1, float s2) {
    const float alt = (n < 2) ? lambda : fmax_ref(s2, lambda);
    return (n == 0) ? 0.f : s1 - alt;
  };
  float c_s[KP][2], c_b[KP][2], c_r[KP][2];
  int c_xr[KP][2], c_n[KP];
  uint32_t c_kk[KP];            // k of candidate 0 | k of candidate 1 << 8
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const int x = tid + k * nt;
    c_n[k] = 0; c_kk[k] = 0u;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      c_s[k][j] = kNone; c_b[k][j] = 0.f; c_r[k][j] = 0.f; c_xr[k][j] = -1;
    }
    if (x < W) {
      const uint4 r = recs[x];
      int n = 0;
      if (r.x) {
        c_s[k][0] = key_score(r.x); c_kk[k] = key_k(r.x); n = 1;
        if (r.y) { c_s[k][1] = key_score(r.y); c_kk[k] |= key_k(r.y) << 8; n = 2; }
      }
      c_n[k] = n;
      g_margin[row + x] = margin_of(n, c_s[k][0], c_s[k][1]);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const uint32_t kj = (c_kk[k] >> (8 * j)) & 0xFFu;
        const int xr = x - (dmin + int(kj));
        if (j < n && xr >= 0 && xr < W) {
          c_xr[k][j] = xr;
          atomicAdd(&S.bstart[xr], 1);
          const unsigned long long key =
              (static_cast<unsigned long long>(ordered(c_s[k][j])) << 32) |
              (uint32_t(0xFFFF - x) << 16) | kj;
          atomicMax(&S.revk[xr], key);
        }
      }
    }
  }
  __syncthreads();

  // --- exclusive scan of the counts (into bstart and the cursors): chunk per
  // thread, inclusive scan within each wave through lane shuffles, then over
  // the wave totals ---
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
    for (int i = lo; i < hi; ++i) {
      const int c = S.bstart[i];
      S.bstart[i] = run;
      S.cursor[i] = run;
      run += c;
    }
    if (tid == nt - 1) S.bstart[W] = S.scan[nwave - 1];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KP; ++k) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int xr = c_xr[k][j];
      if (xr >= 0) S.items[atomicAdd(&S.cursor[xr], 1)] = int16_t(2 * (tid + k * nt) + j);
    }
  }
  __syncthreads();

  // --- message passing ---
  // A pixel's candidates, beta and rho live in its owner's registers across
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
  int* const bti = S.cursor;
  int b_cnt[KP], b_it[KP][kBucketRegs];
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const int xr = tid + k * nt;
    b_cnt[k] = 0;
#pragma unroll
    for (int i = 0; i < kBucketRegs; ++i) b_it[k][i] = 0;
    if (xr < W) {
      const int lo = S.bstart[xr], cnt = S.bstart[xr + 1] - lo;
      b_cnt[k] = cnt;
#pragma unroll
      for (int i = 0; i < kBucketRegs; ++i) if (i < cnt) b_it[k][i] = S.items[lo + i];
    }
  }
  for (int it = 0; it < F.iters; ++it) {
    // rho of every pixel from its own betas. The reference's top-2 (strictly
    // greater displaces, from kNone) over beta 0 then beta 1 excludes to
    // excl(0) = beta 1 (if present and above kNone, else kNone) and
    // excl(1) = beta 0 (if above kNone, else kNone).
#pragma unroll
    for (int k = 0; k < KP; ++k) {
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
    for (int k = 0; k < KP; ++k) {
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
      if (cnt > kBucketRegs) {
        const int lo = S.bstart[xr];
        for (int i = kBucketRegs; i < cnt; ++i) {
          const int idx = S.items[lo + i];
          push(S.rho[idx], idx);
        }
      }
      if (xr < W) {
        S.bt[xr] = make_float2(t1, t2);
        bti[xr] = ti;
      }
    }
    __syncthreads();
    // beta of every bucketed candidate from its bucket's top-2
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int x = tid + k * nt;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int xr = c_xr[k][j];          // -1: absent or outside the row
        const int xq = max(xr, 0);
        const float2 tb = S.bt[xq];
        const float ex = bti[xq] == 2 * x + j ? tb.y : tb.x;
        const float nb = keep * (c_s[k][j] - fmax_ref(gamma, ex)) + damping * c_b[k][j];
        c_b[k][j] = xr >= 0 ? nb : c_b[k][j];
      }
    }
  }

  // --- belief and gates (registers), then the claim: each target right pixel
  // goes to its claimant of highest belief, the smallest x among equals. The
  // claim keys take the rho storage, last read before the final barrier of
  // the iterations (or never, without iterations). ---
  unsigned long long* claim = reinterpret_cast<unsigned long long*>(S.rho);
  for (int x = tid; x < W; x += nt) claim[x] = 0ull;
  float c_best[KP];
  int c_bj[KP];
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const int x = tid + k * nt;
    const int n = c_n[k];
    float bst = kNone;
    int bj = -1;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const float c = c_s[k][j];
      if (j >= n || c <= lambda) continue;
      const float bel = (F.iters > 0) ? (c_b[k][j] + c_r[k][j] - c) : c;
      if (bel > bst) { bst = bel; bj = j; }
    }
    if (bj >= 0 && F.min_margin > 0.f && margin_of(n, c_s[k][0], c_s[k][1]) < F.min_margin)
      bj = -1;
    if (bj >= 0 && F.min_ratio > 0.f) {
      const float s2 = (n > 1) ? fmax_ref(c_s[k][1], lambda) : lambda;
      const float c1 = fmax_ref(1.f - c_s[k][0], 1e-6f);
      if ((1.f - s2) / c1 < F.min_ratio) bj = -1;
    }
    const int kb = int((c_kk[k] >> (bj == 1 ? 8 : 0)) & 0xFFu);
    if (bj >= 0 && x - (dmin + kb) < 0) bj = -1;
    c_best[k] = bst;
    c_bj[k] = (x < W) ? bj : -1;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const int bj = c_bj[k];
    if (bj < 0) continue;
    const int x = tid + k * nt;
    const int xr = x - (dmin + int((c_kk[k] >> (8 * bj)) & 0xFFu));
    const unsigned long long key =
        (static_cast<unsigned long long>(ordered(c_best[k])) << 32) | (0xFFFFFFFFu - uint32_t(x));
    atomicMax(&claim[xr], key);
  }
  __syncthreads();

  // --- outputs ---
  const size_t WH = size_t(W) * F.H;
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const int x = tid + k * nt;
    if (x >= W) continue;
    const int bj = c_bj[k];
    const int n = c_n[k];
    float out = __int_as_float(0x7fc00000);
    if (bj >= 0) {
      const int d = dmin + int((c_kk[k] >> (8 * bj)) & 0xFFu);
      const int xr = x - d;
      const uint32_t mine = 0xFFFFFFFFu - uint32_t(x);
      if (uint32_t(claim[xr] & 0xFFFFFFFFull) == mine) {
        float off = 0.f;
        if (F.subpixel && bj == 0) {
          const uint32_t nb = recs[x].z;
          const float cm = nb_score(nb & 0xFFFFu), cp = nb_score(nb >> 16);
          if (cm > -1e29f && cp > -1e29f) {
            const float c0 = c_s[k][0];
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
      if (xr >= 0 && xr < W) {
        const unsigned long long rk = S.revk[xr];
        if (rk) lrc2 = fabsf(out - float(dmin + int(rk & 0xFFFFu)));
      }
    }
    if (F.cand6) {
      DE_GLOBAL float* c6 = gptr(F.cand6) + row + x;
      c6[0] = c_s[k][0];
      c6[WH] = c_s[k][1];
      c6[2 * WH] = n > 0 ? float(dmin + int(c_kk[k] & 0xFFu)) : kNone;
      c6[3 * WH] = n > 1 ? float(dmin + int(c_kk[k] >> 8)) : kNone;
      c6[4 * WH] = kNone;   // exact reverse match (--lrc) is written by k_lrc_exact
      c6[5 * WH] = lrc2;
    }
    if (F.q4) {
      gptr(F.q4)[row + x] = (fin && out > 0.f) ? int16_t(roundf(out * 16.0f)) : int16_t(-32768);
    }
    if (F.conf) {
      const bool has = fin && out > 0.f;
      gptr(F.conf)[row + x] =
          has ? uint8_t(roundf(confidence_dev(c_s[k][0], c_s[k][1], lambda,
                                              (lrc2 == lrc2) ? lrc2 : 0.f, 1.f) * 255.f))
              : uint8_t(0);
    }
  }
}

size_t solve_lds_bytes(int W) {
  auto r = [](size_t v) { return (v + 15) & ~size_t(15); };
  return r(8 * size_t(W)) * 3 + r(4 * size_t(W + 1)) + r(4 * size_t(W)) * 2 +
         r(4 * size_t(kSolveThreads));
}

// Each instantiation may ask for up to 160 KiB of dynamic LDS (36 B a pixel at
// rows of 4096); the attribute is set once per instantiation.
template <int KP, int NT>
static hipError_t solve_lds_attr() {
  static const hipError_t attr = hipFuncSetAttribute(
      reinterpret_cast<const void*>(&k_solve<KP, NT>), hipFuncAttributeMaxDynamicSharedMemorySize,
      160 * 1024);
  return attr;
}

// maxW: widest row of the launch (at most kMaxW).
hipError_t launch_solve(const BatchView& b, int items, size_t lds, int maxW, hipStream_t s) {
  // Rows of 513..1024 pixels: 1024 threads of one pixel each (at 848x480x128
  // 23.5 us against 26 with 512 threads of two; four frames in flight 0.136
  // against 0.139 ms a frame, two in flight 0.157 against 0.153). Otherwise
  // 512 threads of 1, 4 or 8 pixels.
  const int kp = (maxW + kSolveThreads - 1) / kSolveThreads;
  hipError_t e = hipSuccess;
  if (kp <= 1) {
    if ((e = solve_lds_attr<1, kSolveThreads>()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_solve<1>, dim3(items), dim3(kSolveThreads), lds, s, b);
  } else if (kp == 2) {
    if ((e = solve_lds_attr<1, 2 * kSolveThreads>()) != hipSuccess) return e;
    hipLaunchKernelGGL((k_solve<1, 2 * kSolveThreads>), dim3(items), dim3(2 * kSolveThreads), lds,
                       s, b);
  } else if (kp <= 4) {
    if ((e = solve_lds_attr<4, kSolveThreads>()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_solve<4>, dim3(items), dim3(kSolveThreads), lds, s, b);
  } else if (kp <= 8) {
    if ((e = solve_lds_attr<8, kSolveThreads>()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_solve<8>, dim3(items), dim3(kSolveThreads), lds, s, b);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace de
