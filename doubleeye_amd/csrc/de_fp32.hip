// The reference's FP32 full-volume path on gfx950: the float recursive filter
// (`--dump-vol`, `--no-blockwise`) and the full-volume row solver solve_row
// (`--topk 0`).
//
//   k_f32_coeffs  ax, ay planes from the Q15 table (de_dense.cpp:547-560)
//   k_f32_row     one wave per (row, 64 disparities): float score, H-fwd and
//                 H-bwd (de_dense.cpp:379-396), lanes = disparities
//   k_f32_col     one wave per (column, 64 disparities): V-fwd and V-bwd
//                 (de_dense.cpp:510-527), then the volume [y][x][k] with -1e30
//                 outside the valid window (de_dense.cpp:1404-1411)
//   k_f32_solve   one workgroup per image row: solve_row (de_dense.cpp:566-712):
//                 margins, `iters` rounds of rho (a wave per left pixel, lanes
//                 over its disparities) and beta (a wave per right pixel, lanes
//                 over its claimants on the diagonal x = xr + dmin + k), best
//                 belief, greedy claim as a per-right-pixel argmax (LDS 64-bit
//                 atomic max; equal beliefs: the smaller x, where the reference
//                 follows std::sort's unstable order), parabola sub-pixel.
// Top-2 with "first index among equals" is a wave max over the packed key
// (value, ~k); the runner-up is a second max with that lane's entry removed,
// which is exactly what the sequential insertion leaves. Compiled with
// -ffp-contract=off: every float operation rounds as the reference's does.
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

constexpr int kSolveT = 1024;
constexpr int kSolveWaves = kSolveT / kWave;
constexpr int kMaxChunks = kMaxD / kWave;   // D <= 256: up to 4 disparities per lane
constexpr float kRfScale = 1.f / 32768.f;

struct F32Desc {
  const uint8_t* L;
  const uint64_t* cl;
  const uint64_t* cr;
  const uint16_t* rf;
  uint16_t* ax;
  uint16_t* ay;
  float* work;       // [y][x][Dpad]
  float* vol;        // [y][x][D]
  float* beta;       // [y][x][D]
  float* rho;        // [y][x][D]
  float* disp;
  float* margin;
  int W, H, D, Dpad, G, dmin, hw, hh, iters, subpixel;
  float half, inv, lambda, gamma, damping, min_margin;
};

__device__ __forceinline__ uint32_t ordered(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u << 1) == 0u) u = 0u;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unordered(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}
__device__ __forceinline__ float fmax_ref(float a, float b) { return (a < b) ? b : a; }

// Sequential top-2 over values v[c] at k = c*64 + lane (valid[c]), in
// ascending k, started from (kNone, kNone, -1): b1 = the largest value above
// kNone, i1 its first index (-1 if none), b2 = the largest other value, at
// least kNone.
struct WTop2 {
  float b1, b2;
  int i1;
};
// Largest u32 over the wave (0: none) by DPP: row shifts 1, 2, 4, 8 leave each
// row's maximum in its lane 15, two row broadcasts carry them to lane 63.
// Every lane must be active.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  v = max(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x111, 0xF, 0xF, false)));
  v = max(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x112, 0xF, 0xF, false)));
  v = max(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x114, 0xF, 0xF, false)));
  v = max(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x118, 0xF, 0xF, false)));
  v = max(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x142, 0xA, 0xF, false)));
  v = max(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x143, 0xC, 0xF, false)));
  return __builtin_amdgcn_readlane(v, 63);
}

// The smallest k = c*64 + lane whose key is m (-1: none), by ballots in
// ascending chunk order. Every lane must be active.
__device__ __forceinline__ int wave_first(const uint32_t* key, int nchunk, uint32_t m) {
  for (int c = 0; c < nchunk; ++c) {
    const uint64_t b = __ballot(key[c] == m);
    if (b) return c * kWave + __ffsll(static_cast<unsigned long long>(b)) - 1;
  }
  return -1;
}

// The same order as packed keys (ordered value << 32 | ~k): the largest
// value, the smallest k among equal values.
__device__ WTop2 wave_top2(const float* v, const bool* valid, int nchunk, int lane) {
  uint32_t key[kMaxChunks];
  uint32_t lm = 0u;
  for (int c = 0; c < nchunk; ++c) {
    key[c] = valid[c] ? ordered(v[c]) : 0u;
    lm = max(lm, key[c]);
  }
  WTop2 t;
  t.b1 = kNone;
  t.b2 = kNone;
  t.i1 = -1;
  const uint32_t m = wave_max_u32(lm);
  if (m == 0u) return t;
  const float b1 = unordered(m);
  if (!(b1 > kNone)) return t;
  t.b1 = b1;
  t.i1 = wave_first(key, nchunk, m);
  uint32_t lm2 = 0u;
  for (int c = 0; c < nchunk; ++c)
    if (c * kWave + lane != t.i1) lm2 = max(lm2, key[c]);
  const uint32_t m2 = wave_max_u32(lm2);
  if (m2 != 0u) {
    const float s2 = unordered(m2);
    t.b2 = s2 > kNone ? s2 : kNone;
  }
  return t;
}

// A pixel's best belief key over its disparities: (ordered(bel) << 32 |
// ~k) of the largest belief, smallest k among equals; 0 when none.
__device__ __forceinline__ uint64_t wave_best_key(const uint32_t* key, int nchunk) {
  uint32_t lm = 0u;
  for (int c = 0; c < nchunk; ++c) lm = max(lm, key[c]);
  const uint32_t m = wave_max_u32(lm);
  if (m == 0u) return 0ull;
  const int k = wave_first(key, nchunk, m);
  return (uint64_t(m) << 32) | uint64_t(0xFFFFFFFFu - uint32_t(k));
}

__global__ void k_f32_coeffs(F32Desc f) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= f.W * f.H) return;
  const int x = i % f.W, y = i / f.W;
  const int l = f.L[i];
  f.ax[i] = x ? f.rf[abs(l - int(f.L[i - 1]))] : f.rf[0];
  f.ay[i] = y ? f.rf[abs(l - int(f.L[i - f.W]))] : f.rf[0];
}

// The loops take their operands 8 steps at a time before the dependent
// chain and its in-place stores (the compiler cannot move a load above a store
// to the work plane, which it cannot prove disjoint), so 8 loads are in flight.
constexpr int kF32Batch = 8;

__global__ __launch_bounds__(256) void k_f32_row(F32Desc f) {
  const int item = __builtin_amdgcn_readfirstlane(int(blockIdx.x) * 4 + int(threadIdx.x >> 6));
  if (item >= f.H * f.G) return;
  const int lane = threadIdx.x & 63;
  const int y = item / f.G, g = item % f.G;
  const int k = g * kWave + lane;
  const int d = f.dmin + k;
  const bool live = k < f.D && y >= f.hh && y < f.H - f.hh;
  const size_t row = size_t(y) * f.W;
  const int W = f.W;
  float fv = 0.f;
  for (int x0 = 0; x0 < W; x0 += kF32Batch) {
    float as[kF32Batch], P[kF32Batch];
#pragma unroll
    for (int u = 0; u < kF32Batch; ++u) {
      const int x = x0 + u;
      as[u] = 0.f;
      P[u] = 0.f;
      if (x < W) {
        const uint64_t a = f.cl[row + x];
        as[u] = float(f.ax[row + x]) * kRfScale;
        if (live && x >= f.hw + d && x < W - f.hw)
          P[u] = (f.half - float(__popcll(a ^ f.cr[row + x - d]))) * f.inv;
      }
    }
#pragma unroll
    for (int u = 0; u < kF32Batch; ++u) {
      const int x = x0 + u;
      if (x < W) {
        fv = x == 0 ? P[u] : P[u] + as[u] * (fv - P[u]);
        f.work[(row + x) * f.Dpad + k] = fv;
      }
    }
  }
  for (int x1 = W - 2; x1 >= 0; x1 -= kF32Batch) {       // columns x1, x1-1, ..., x1-7
    float F[kF32Batch], as[kF32Batch];
#pragma unroll
    for (int u = 0; u < kF32Batch; ++u) {
      const int x = x1 - u;
      F[u] = 0.f;
      as[u] = 0.f;
      if (x >= 0) {
        F[u] = f.work[(row + x) * f.Dpad + k];
        as[u] = float(f.ax[row + x + 1]) * kRfScale;
      }
    }
#pragma unroll
    for (int u = 0; u < kF32Batch; ++u) {
      const int x = x1 - u;
      if (x >= 0) {
        fv = F[u] + as[u] * (fv - F[u]);
        f.work[(row + x) * f.Dpad + k] = fv;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_f32_col(F32Desc f) {
  const int item = __builtin_amdgcn_readfirstlane(int(blockIdx.x) * 4 + int(threadIdx.x >> 6));
  if (item >= f.W * f.G) return;
  const int lane = threadIdx.x & 63;
  const int x = item / f.G, g = item % f.G;
  const int k = g * kWave + lane;
  const int d = f.dmin + k;
  const size_t W = size_t(f.W);
  const int H = f.H;
  float prev = f.work[size_t(x) * f.Dpad + k];
  for (int y0 = 1; y0 < H; y0 += kF32Batch) {
    float c[kF32Batch], as[kF32Batch];
#pragma unroll
    for (int u = 0; u < kF32Batch; ++u) {
      const int y = y0 + u;
      c[u] = 0.f;
      as[u] = 0.f;
      if (y < H) {
        const size_t i = size_t(y) * W + x;
        c[u] = f.work[i * f.Dpad + k];
        as[u] = float(f.ay[i]) * kRfScale;
      }
    }
#pragma unroll
    for (int u = 0; u < kF32Batch; ++u) {
      const int y = y0 + u;
      if (y < H) {
        prev = c[u] + as[u] * (prev - c[u]);
        f.work[(size_t(y) * W + x) * f.Dpad + k] = prev;
      }
    }
  }
  const bool xok = k < f.D && x >= f.hw + d && x < f.W - f.hw;
  auto emit = [&](int y, float v) {
    if (k < f.D)
      f.vol[(size_t(y) * W + x) * f.D + k] = (xok && y >= f.hh && y < f.H - f.hh) ? v : kNone;
  };
  float nxt = prev;
  emit(H - 1, nxt);
  for (int y1 = H - 2; y1 >= 0; y1 -= kF32Batch) {       // rows y1, y1-1, ..., y1-7
    float c[kF32Batch], as[kF32Batch];
#pragma unroll
    for (int u = 0; u < kF32Batch; ++u) {
      const int y = y1 - u;
      c[u] = 0.f;
      as[u] = 0.f;
      if (y >= 0) {
        const size_t i = size_t(y) * W + x;
        c[u] = f.work[i * f.Dpad + k];
        as[u] = float(f.ay[i + W]) * kRfScale;
      }
    }
#pragma unroll
    for (int u = 0; u < kF32Batch; ++u) {
      const int y = y1 - u;
      if (y >= 0) {
        nxt = c[u] + as[u] * (nxt - c[u]);
        emit(y, nxt);
      }
    }
  }
}

// The solver's window of valid k for left pixel x (the reference passes the
// census border 3 here whatever the window, de_dense.cpp:1541).
__device__ __forceinline__ void window(int x, int W, int D, int dmin, int* lo, int* hi) {
  const int hw = 3;
  int a = 0, b = 0;
  if (x >= hw && x < W - hw) {
    a = max(0, x - (W - hw) + 1 - dmin + 1);
    b = min(D, x - hw - dmin + 1);
    if (b < a) b = a;
  }
  *lo = a;
  *hi = b;
}

// The beta half-step's top-2 over a right pixel's claimants (x = xr + dmin + k)
// from its packed keys (ordered value << 32 | ~k): the best key, and the best
// key with another k. The same rules as wave_top2.
struct Diag {
  float b1, b2;
  int i1;
};
__device__ __forceinline__ Diag diag_top2(uint64_t top, uint64_t sec) {
  Diag t{kNone, kNone, -1};
  if (top == 0) return t;
  const float m = unordered(uint32_t(top >> 32));
  if (!(m > kNone)) return t;
  t.b1 = m;
  t.i1 = int(0xFFFFFFFFu - uint32_t(top));
  if (sec != 0) {
    const float s2 = unordered(uint32_t(sec >> 32));
    t.b2 = s2 > kNone ? s2 : kNone;
  }
  return t;
}

__device__ __forceinline__ uint64_t f32_key(float v, int k) {
  return (uint64_t(ordered(v)) << 32) | uint64_t(0xFFFFFFFFu - uint32_t(k));
}

// Every pass walks the row's [x][k] planes with lanes over k (coalesced). The
// beta half-step's top-2 along each right pixel's diagonal is built with LDS
// 64-bit atomic maxima of packed keys in the rho pass that writes the rho
// values: the best, and the runner-up (the best with another k) from the
// smaller of each key and the key it displaced. Keys are unique in a diagonal
// and order by value then smaller k, as wave_top2's do, so the order of the
// atomics does not matter. Two key sets alternate by iteration, so that the
// beta update of iteration it - 1 runs at the head of iteration it's rho pass
// (a pixel's own elements, before their top-2 over k).
// NC: disparity chunks of 64 (ceil(D / 64)), a template so the chunk loops unroll.
template <int NC>
__global__ __launch_bounds__(kSolveT) void k_f32_solve(F32Desc f) {
  extern __shared__ unsigned long long claim[];   // [4W]: key sets; claim[W] at the end
  const int y = blockIdx.x;
  const int W = f.W, D = f.D, dmin = f.dmin;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const size_t rb = size_t(y) * W * D;
  const float* __restrict__ vol = f.vol + rb;
  float* __restrict__ beta = f.beta + rb;
  float* __restrict__ rho = f.rho + rb;
  constexpr int nchunk = NC;
  const float keep = 1.f - f.damping;
  // Messages start at zero. With iterations, nothing is filled: iteration 0
  // takes its old beta and rho as the literal 0.0f a fill would have stored
  // (the same operands, so the same bits), and every element of a pixel's
  // window is written before it is read again (the window keeps x - dmin - k
  // >= 3, so each has a diagonal).
  if (f.iters == 0)
    for (int t = threadIdx.x; t < W * D; t += kSolveT) {
      beta[t] = 0.f;
      rho[t] = 0.f;
    }
  auto dk = [&](int set, int which) { return claim + size_t(set * 2 + which) * W; };
  for (int t = threadIdx.x; t < 2 * W; t += kSolveT) claim[t] = 0ull;   // set 0
  // margins (scores above -1e29 only)
  for (int x = wv; x < W; x += kSolveWaves) {
    int lo, hi;
    window(x, W, D, dmin, &lo, &hi);
    float v[kMaxChunks];
    bool ok[kMaxChunks];
    #pragma unroll
      for (int c = 0; c < nchunk; ++c) {
      const int k = c * kWave + lane;
      ok[c] = k >= lo && k < hi;
      v[c] = ok[c] ? vol[size_t(x) * D + k] : 0.f;
      ok[c] = ok[c] && v[c] > -1e29f;
    }
    const WTop2 t = wave_top2(v, ok, nchunk, lane);
    if (lane == 0) {
      const float alt = (t.b2 < -1e29f) ? f.lambda : fmax_ref(t.b2, f.lambda);
      f.margin[size_t(y) * W + x] = (t.i1 < 0) ? 0.f : t.b1 - alt;
    }
  }
  __syncthreads();
  // beta(x, k) from its diagonal's top-2 in key set `set` (the beta half-step),
  // given the score s and the old beta b0; returns the new value
  auto beta_new = [&](int set, int x, int k, float s, float b0) {
    const int xr = x - dmin - k;
    const Diag t = diag_top2(dk(set, 0)[xr], dk(set, 1)[xr]);
    const float ex = (k == t.i1) ? t.b2 : t.b1;
    return keep * (s - fmax_ref(f.gamma, ex)) + f.damping * b0;
  };
  for (int it = 0; it < f.iters; ++it) {
    const int cs = it & 1, ps = cs ^ 1;
    // beta of iteration it - 1, then rho: a left pixel's disparities compete;
    // each new rho enters its diagonal's best key
    for (int x = wv; x < W; x += kSolveWaves) {
      int lo, hi;
      window(x, W, D, dmin, &lo, &hi);
      if (hi <= lo) continue;
      float v[kMaxChunks], sv[kMaxChunks], r0[kMaxChunks];
      bool ok[kMaxChunks];
      // every operand of the pixel's elements is loaded once
      #pragma unroll
      for (int c = 0; c < nchunk; ++c) {
        const int k = c * kWave + lane;
        ok[c] = k >= lo && k < hi;
        const size_t i = size_t(x) * D + k;
        sv[c] = ok[c] ? vol[i] : 0.f;
        // beta is first written at iteration 1, rho at iteration 0
        v[c] = (ok[c] && it > 1) ? beta[i] : 0.f;
        r0[c] = (ok[c] && it > 0) ? rho[i] : 0.f;
      }
      if (it > 0) {
        #pragma unroll
        for (int c = 0; c < nchunk; ++c) {
          if (!ok[c]) continue;
          const int k = c * kWave + lane;
          v[c] = beta_new(ps, x, k, sv[c], v[c]);
          beta[size_t(x) * D + k] = v[c];
        }
      }
      const WTop2 t = wave_top2(v, ok, nchunk, lane);
      #pragma unroll
      for (int c = 0; c < nchunk; ++c) {
        if (!ok[c]) continue;
        const int k = c * kWave + lane;
        const size_t i = size_t(x) * D + k;
        const float ex = (k == t.i1) ? t.b2 : t.b1;
        const float tgt = sv[c] - fmax_ref(f.lambda, ex);
        const float r = keep * tgt + f.damping * r0[c];
        rho[i] = r;
        const int xr = x - dmin - k;
        if (xr >= 0) {
          // best key; the smaller of this key and the one it displaced goes to
          // the runner-up, so that every key but the final best ends there
          const uint64_t key = f32_key(r, k);
          const uint64_t old = atomicMax(&dk(cs, 0)[xr], key);
          atomicMax(&dk(cs, 1)[xr], key > old ? old : key);
        }
      }
    }
    __syncthreads();
    // the other key set is cleared for the next iteration (its last reads
    // were in the pass above)
    for (int t = threadIdx.x; t < 2 * W; t += kSolveT) dk(ps, 0)[t] = 0ull;
    __syncthreads();
  }
  // the last iteration's beta half-step (old beta: zero after one iteration)
  if (f.iters > 0) {
    const int ls = (f.iters - 1) & 1;
    for (int x = wv; x < W; x += kSolveWaves) {
      int lo, hi;
      window(x, W, D, dmin, &lo, &hi);
      #pragma unroll
      for (int c = 0; c < nchunk; ++c) {
        const int k = c * kWave + lane;
        if (k < lo || k >= hi) continue;
        const size_t i = size_t(x) * D + k;
        beta[i] = beta_new(ls, x, k, vol[i], f.iters > 1 ? beta[i] : 0.f);
      }
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < W; t += kSolveT) claim[t] = 0ull;
  __syncthreads();
  // best belief per left pixel, then the claim of its right pixel
  for (int x = wv; x < W; x += kSolveWaves) {
    int lo, hi;
    window(x, W, D, dmin, &lo, &hi);
    uint32_t bk[kMaxChunks];
    #pragma unroll
      for (int c = 0; c < nchunk; ++c) {
      const int k = c * kWave + lane;
      bk[c] = 0u;
      if (k < lo || k >= hi) continue;
      const size_t i = size_t(x) * D + k;
      const float s = vol[i];
      if (s <= f.lambda) continue;
      const float bel = beta[i] + rho[i] - s;
      if (!(bel > -1e30f)) continue;
      bk[c] = ordered(bel);
    }
    const uint64_t top = wave_best_key(bk, nchunk);
    if (lane == 0) {
      f.disp[size_t(y) * W + x] = NAN;
      if (top != 0) {
        const int k = int(0xFFFFFFFFu - uint32_t(top));
        const int xr = x - (dmin + k);
        const bool gated = f.min_margin > 0.f && f.margin[size_t(y) * W + x] < f.min_margin;
        if (!gated && xr >= 0)
          atomicMax(&claim[xr], (top & 0xFFFFFFFF00000000ull) | uint64_t(0xFFFFFFFFu - uint32_t(x)));
      }
    }
  }
  __syncthreads();
  for (int x = wv; x < W; x += kSolveWaves) {
    int lo, hi;
    window(x, W, D, dmin, &lo, &hi);
    uint32_t bk[kMaxChunks];
    #pragma unroll
      for (int c = 0; c < nchunk; ++c) {
      const int k = c * kWave + lane;
      bk[c] = 0u;
      if (k < lo || k >= hi) continue;
      const size_t i = size_t(x) * D + k;
      const float s = vol[i];
      if (s <= f.lambda) continue;
      const float bel = beta[i] + rho[i] - s;
      if (!(bel > -1e30f)) continue;
      bk[c] = ordered(bel);
    }
    const uint64_t top = wave_best_key(bk, nchunk);
    if (lane == 0 && top != 0) {
      const int k = int(0xFFFFFFFFu - uint32_t(top));
      const int d = dmin + k;
      const int xr = x - d;
      if (xr >= 0 &&
          claim[xr] == ((top & 0xFFFFFFFF00000000ull) | uint64_t(0xFFFFFFFFu - uint32_t(x)))) {
        float off = 0.f;
        const float* sp = vol + size_t(x) * D;
        if (f.subpixel && k > lo && k + 1 < hi) {
          const float cm = sp[k - 1], c0 = sp[k], cp = sp[k + 1];
          const float den = 2.f * c0 - cm - cp;
          if (den > 1e-6f) {
            off = 0.5f * (cp - cm) / den;
            if (off > 0.5f) off = 0.5f;
            if (off < -0.5f) off = -0.5f;
          }
        }
        f.disp[size_t(y) * W + x] = float(d) + off;
      }
    }
  }
}

size_t a256(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

// FP32 volume (and optionally the full-volume solve) of one pair whose
// images are on the device. vol_out / disp / margin are device pointers.
static int f32_run(Ctx& x, const uint8_t* dL, const uint8_t* dR, int W, int H, const de_cfg& c,
                   float* vol_out, float* disp, float* margin, bool solve, hipStream_t s) {
  const int D = c.dmax - c.dmin + 1;
  const int G = (D + kWave - 1) / kWave, Dpad = G * kWave;
  const size_t WH = size_t(W) * H;
  const size_t bytes = a256(WH * 8) * 2 + a256(WH * 2) * 2 + a256(WH * Dpad * 4) +
                       (vol_out ? 0 : a256(WH * D * 4)) + (solve ? a256(WH * D * 4) * 2 : 0);
  if (int rc = x.f32_scratch.ensure(bytes)) return rc;
  const uint16_t *rf, *sc;
  int shift;
  if (int rc = x.tables_for(c, &rf, &sc, &shift)) return rc;
  uint8_t* p = static_cast<uint8_t*>(x.f32_scratch.ptr);
  auto take = [&](size_t b) { void* r = p; p += a256(b); return r; };
  F32Desc f;
  std::memset(&f, 0, sizeof(f));
  f.L = dL;
  f.cl = static_cast<uint64_t*>(take(WH * 8));
  f.cr = static_cast<uint64_t*>(take(WH * 8));
  f.ax = static_cast<uint16_t*>(take(WH * 2));
  f.ay = static_cast<uint16_t*>(take(WH * 2));
  f.work = static_cast<float*>(take(WH * Dpad * 4));
  f.vol = vol_out ? vol_out : static_cast<float*>(take(WH * D * 4));
  if (solve) {
    f.beta = static_cast<float*>(take(WH * D * 4));
    f.rho = static_cast<float*>(take(WH * D * 4));
  }
  f.rf = rf;
  f.disp = disp;
  f.margin = margin;
  f.W = W;
  f.H = H;
  f.D = D;
  f.Dpad = Dpad;
  f.G = G;
  f.dmin = c.dmin;
  f.hw = c.census_hw;
  f.hh = c.census_hh;
  f.iters = c.iters;
  f.subpixel = c.subpixel;
  const int nb = (2 * c.census_hw + 1) * (2 * c.census_hh + 1) - 1;
  f.half = float(nb / 2);
  f.inv = 1.f / f.half;
  f.lambda = c.lambda;
  f.gamma = c.gamma;
  f.damping = c.damping;
  f.min_margin = c.min_margin;
  DE_HIP(launch_census_plane(dL, W, H, c.census_hw, c.census_hh, const_cast<uint64_t*>(f.cl), s));
  DE_HIP(launch_census_plane(dR, W, H, c.census_hw, c.census_hh, const_cast<uint64_t*>(f.cr), s));
  hipLaunchKernelGGL(k_f32_coeffs, dim3(unsigned((WH + 255) / 256)), dim3(256), 0, s, f);
  hipLaunchKernelGGL(k_f32_row, dim3(unsigned((H * G + 3) / 4)), dim3(256), 0, s, f);
  hipLaunchKernelGGL(k_f32_col, dim3(unsigned((W * G + 3) / 4)), dim3(256), 0, s, f);
  if (solve) {
    // 32 bytes of LDS per pixel of the row: above the default dynamic limit from W = 2049
    const int nc = (D + kWave - 1) / kWave;
    auto go = [&](auto kern) -> hipError_t {
      const hipError_t a = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               160 * 1024);
      if (a != hipSuccess) return a;
      hipLaunchKernelGGL(kern, dim3(H), dim3(kSolveT), size_t(W) * 32, s, f);
      return hipGetLastError();
    };
    hipError_t e;
    if (nc == 1) e = go(k_f32_solve<1>);
    else if (nc == 2) e = go(k_f32_solve<2>);
    else if (nc == 3) e = go(k_f32_solve<3>);
    else e = go(k_f32_solve<4>);
    DE_HIP(e);
  }
  DE_HIP(hipGetLastError());
  return DE_OK;
}

static int f32_validate(const de_cfg& c, int W, int H) {
  if (int rc = validate(c, W, H)) return rc;
  if (size_t(W) * 32 > 160 * 1024) return fail(DE_ERR_ARG, "W=%d too wide for the row solver", W);
  return DE_OK;
}

}  // namespace de

using namespace de;

namespace {

int f32_host(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H, const de_cfg* cfg,
             float* vol, float* disp, float* margin, bool solve) {
  if (!c || !L || !R || !cfg) return fail(DE_ERR_ARG, "fp32: null argument");
  de_cfg k = *cfg;
  k.lrc = 0;   // lrc is a plane of the blockwise path; the full-volume path has none
  if (int rc = f32_validate(k, W, H)) return rc;
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  const size_t WH = size_t(W) * H;
  const int D = k.dmax - k.dmin + 1;
  const size_t io = a256(WH) * 2 + a256(WH * D * 4) + a256(WH * 4) * 2;
  if (int rc = x.f32_io.ensure(io)) return rc;
  uint8_t* p = static_cast<uint8_t*>(x.f32_io.ptr);
  uint8_t* dL = p;
  uint8_t* dR = p + a256(WH);
  float* dvol = reinterpret_cast<float*>(p + a256(WH) * 2);
  float* ddisp = reinterpret_cast<float*>(p + a256(WH) * 2 + a256(WH * D * 4));
  float* dmarg = ddisp + a256(WH * 4) / 4;
  DE_HIP(hipMemcpyAsync(dL, L, WH, hipMemcpyHostToDevice, x.stream));
  DE_HIP(hipMemcpyAsync(dR, R, WH, hipMemcpyHostToDevice, x.stream));
  if (int rc = f32_run(x, dL, dR, W, H, k, dvol, ddisp, dmarg, solve, x.stream)) return rc;
  if (vol) DE_HIP(hipMemcpyAsync(vol, dvol, WH * D * 4, hipMemcpyDeviceToHost, x.stream));
  if (solve) {
    if (disp) DE_HIP(hipMemcpyAsync(disp, ddisp, WH * 4, hipMemcpyDeviceToHost, x.stream));
    if (margin) DE_HIP(hipMemcpyAsync(margin, dmarg, WH * 4, hipMemcpyDeviceToHost, x.stream));
  }
  DE_HIP(hipStreamSynchronize(x.stream));
  return DE_OK;
}

}  // namespace

extern "C" {

int de_dense_volume_f32(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                        const de_cfg* cfg, float* vol) {
  if (!vol) return fail(DE_ERR_ARG, "de_dense_volume_f32: null output");
  return f32_host(c, L, R, W, H, cfg, vol, nullptr, nullptr, false);
}

int de_dense_run_full(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                      const de_cfg* cfg, float* disp, float* margin) {
  if (!disp || !margin) return fail(DE_ERR_ARG, "de_dense_run_full: null output");
  return f32_host(c, L, R, W, H, cfg, nullptr, disp, margin, true);
}

int de_dense_run_full_dev(de_ctx* c, const uint8_t* dL, const uint8_t* dR, int W, int H,
                          const de_cfg* cfg, float* d_disp, float* d_margin, void* stream) {
  if (!c || !dL || !dR || !cfg || !d_disp || !d_margin)
    return fail(DE_ERR_ARG, "de_dense_run_full_dev: null argument");
  de_cfg k = *cfg;
  k.lrc = 0;
  if (int rc = f32_validate(k, W, H)) return rc;
  std::lock_guard<std::mutex> lk(c->c.mu);
  DE_HIP(hipSetDevice(c->c.device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->c.stream;
  return f32_run(c->c, dL, dR, W, H, k, nullptr, d_disp, d_margin, true, s);
}

}  // extern "C"
