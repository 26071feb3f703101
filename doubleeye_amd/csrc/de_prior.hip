// The prior-masked dense sweep (`de_dense --prior`, `--c2f`) on the MI355X.
//
// --prior (core/tools/de_dense.cpp:797-853, 1156-1296, 1771-1808): the prior
// decides which pixels of which absolute-disparity planes are evaluated; the
// masked kernels live with the full sweep in de_dense.hip. The solver is the
// full sweep's. Sub-pixel is off under a mask (de_dense.cpp:1824-1840).
//
// --c2f (de_dense.cpp:1889-1966): a half-resolution pass (2x2 box average with
// rounding, D/2 planes, one iteration, min_margin 0) gives the prior: holes are
// filled from row neighbours (forward then backward), then column neighbours,
// and the result is upsampled to full resolution (nearest, disparity doubled).
// If holes remain the full sweep runs instead. Every stage runs on the device;
// the host reads one hole count to choose between the two sweeps.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "de_api_internal.h"

namespace de {

namespace {

size_t a256(size_t v) { return (v + 255) & ~size_t(255); }

__global__ __launch_bounds__(256) void k_half(const uint8_t* __restrict__ L,
                                              const uint8_t* __restrict__ R, int W, int W2,
                                              int H2, uint8_t* __restrict__ L2,
                                              uint8_t* __restrict__ R2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= W2 * H2) return;
  const int y = i / W2, x = i - y * W2;
  const size_t a = size_t(2 * y) * W + 2 * x;
  L2[i] = uint8_t((int(L[a]) + L[a + 1] + L[a + W] + L[a + W + 1] + 2) / 4);
  R2[i] = uint8_t((int(R[a]) + R[a + 1] + R[a + W] + R[a + W + 1] + 2) / 4);
}

// Nearest-neighbour hole fill of n values at stride `st`, one wave per line:
// the reference's forward carry then backward carry leaves every hole with the
// last finite value before it, or, for the holes before the first finite value,
// that first value. The forward carry is a prefix max of finite indices;
// finite values are only read and holes only written, so the line is filled in
// place. Returns (through the wave) the holes left.
__device__ int fill_line(float* d, int n, size_t st) {
  const int lane = __lane_id();
  int first = n;
  for (int b = 0; b < n; b += kWave) {
    const int i = b + lane;
    const bool f = i < n && isfinite(d[size_t(i) * st]);
    const uint64_t m = __ballot(f);
    if (m) {
      first = b + __ffsll(static_cast<unsigned long long>(m)) - 1;
      break;
    }
  }
  if (first == n) return n;   // nothing to carry: all holes stay
  const float f0 = d[size_t(first) * st];
  int carry = -1;
  for (int b = 0; b < n; b += kWave) {
    const int i = b + lane;
    const bool in = i < n;
    const float v = in ? d[size_t(i) * st] : 0.f;
    const bool fin = in && isfinite(v);
    int idx = fin ? i : -1;
    for (int o = 1; o < kWave; o <<= 1) {
      const int t = __shfl_up(idx, o);
      if (lane >= o) idx = max(idx, t);
    }
    idx = max(idx, carry);
    if (in && !fin) d[size_t(i) * st] = idx >= 0 ? d[size_t(idx) * st] : f0;
    carry = __shfl(idx, kWave - 1);
  }
  return 0;
}

__global__ __launch_bounds__(256) void k_fill_rows(float* __restrict__ d, int W, int H) {
  const int y = blockIdx.x * (256 / kWave) + int(threadIdx.x) / kWave;
  if (y >= H) return;   // uniform over the wave
  (void)fill_line(d + size_t(y) * W, W, 1);
}

__global__ __launch_bounds__(256) void k_fill_cols(float* __restrict__ d, int W, int H,
                                                   int* __restrict__ holes) {
  const int x = blockIdx.x * (256 / kWave) + int(threadIdx.x) / kWave;
  if (x >= W) return;   // uniform over the wave
  const int left = fill_line(d + x, H, size_t(W));
  if (left && __lane_id() == 0) atomicAdd(holes, left);
}

// Nearest upsample of a half-resolution disparity map, doubled.
__global__ __launch_bounds__(256) void k_upsample(const float* __restrict__ half, int W, int H,
                                                  float* __restrict__ full) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= W * H) return;
  const int W2 = W / 2, H2 = H / 2;
  const int y = i / W, x = i - y * W;
  full[i] = 2.f * half[size_t(min(H2 - 1, y / 2)) * W2 + min(W2 - 1, x / 2)];
}

__global__ void k_zero_int(int* p) { *p = 0; }

size_t mask_bytes(int W, int H, int D) {
  return a256(size_t(W) * H * 4) + 2 * a256(size_t(H) * D * 4) + 2 * a256(size_t(D) * 16);
}

MaskBufs carve_mask(void* base, int W, int H, int D) {
  uint8_t* p = static_cast<uint8_t*>(base);
  auto take = [&](size_t bytes) { uint8_t* q = p; p += a256(bytes); return q; };
  MaskBufs m;
  m.dc = reinterpret_cast<int*>(take(size_t(W) * H * 4));
  m.mx0 = reinterpret_cast<int*>(take(size_t(H) * D * 4));
  m.mx1 = reinterpret_cast<int*>(take(size_t(H) * D * 4));
  m.box = reinterpret_cast<int4*>(take(size_t(D) * 16));
  m.rect = reinterpret_cast<int4*>(take(size_t(D) * 16));
  return m;
}

}  // namespace

// The masked sweep of one frame; inputs and outputs are device pointers, the
// prior a full-resolution device plane without holes.
static int prior_dev(Ctx& x, const uint8_t* dL, const uint8_t* dR, int W, int H, de_cfg c,
                     const float* d_prior, int band, float* disp, float* margin, float* cand6,
                     hipStream_t s) {
  c.subpixel = 0;
  if (int rc = validate(c, W, H)) return rc;
  const int D = c.dmax - c.dmin + 1;
  if (int rc = x.scratch.ensure(scratch_bytes(W, H, D))) return rc;
  if (int rc = x.pr_buf.ensure(mask_bytes(W, H, D))) return rc;
  const Scratch sp = carve_scratch(x.scratch.ptr, W, H, D);
  const MaskBufs m = carve_mask(x.pr_buf.ptr, W, H, D);
  FrameDesc fd;
  if (int rc = x.frame_desc(c, W, H, dL, dR, disp, margin, cand6, nullptr, nullptr, sp, &fd))
    return rc;
  Ctx::Prepared P;
  if (int rc = x.prepare(&fd, 1, &P)) return rc;
  const size_t bytes = P.desc_bytes();
  Ctx::Slot* slot;
  if (int rc = x.acquire_slot(bytes, &slot)) return rc;
  P.pack(slot->host);
  DE_HIP(hipMemcpyAsync(slot->dev, slot->host, bytes, hipMemcpyHostToDevice, s));
  BatchView bv;
  bv.frames = static_cast<const FrameDesc*>(slot->dev);
  bv.start = reinterpret_cast<const int*>(static_cast<uint8_t*>(slot->dev) + sizeof(FrameDesc));
  bv.n = 1;
  auto items = [&](int st) { return P.start[size_t(st) * 2 + 1]; };
  hipError_t e;
  if ((e = launch_prep(bv, items(kStagePrep), fd.hw, fd.hh, s)) != hipSuccess)
    return hip_fail(e, "k_prep");
  if ((e = launch_masked(P.frames[0], bv.frames, d_prior, band, m, P.lds_row, fd.lrc != 0, s)) !=
      hipSuccess)
    return hip_fail(e, "masked sweep");
  if ((e = launch_solve(bv, items(kStageSolve), P.lds_solve, P.max_w, s)) != hipSuccess)
    return hip_fail(e, "k_solve");
  if (fd.lrc && (e = launch_lrc_plane(bv, items(kStageSolve), s)) != hipSuccess)
    return hip_fail(e, "k_lrc_plane");
  DE_HIP(hipEventRecord(slot->done, s));
  slot->used = true;
  return DE_OK;
}

// --c2f on device pointers. *coarse gets the device address of the filled
// W/2 x H/2 coarse disparity (valid until the next call on the context);
// *used_mask whether the masked sweep ran.
static int c2f_dev(Ctx& x, const uint8_t* dL, const uint8_t* dR, int W, int H, de_cfg c,
                   int band, float* disp, float* margin, float* cand6, const float** coarse,
                   int* used_mask, hipStream_t s) {
  c.subpixel = 0;
  if (W < 2 || H < 2) return fail(DE_ERR_ARG, "c2f: image %dx%d has no half level", W, H);
  if (int rc = validate(c, W, H)) return rc;
  const int D = c.dmax - c.dmin + 1;
  const int W2 = W / 2, H2 = H / 2;
  de_cfg cc = c;
  cc.dmin = std::max(1, c.dmin / 2);
  cc.dmax = (c.dmax + 1) / 2;
  cc.min_margin = 0.f;
  cc.iters = 1;
  if (int rc = validate(cc, W2, H2)) return rc;
  // the fine level's scratch first: the coarse one fits in it
  if (int rc = x.scratch.ensure(scratch_bytes(W, H, D))) return rc;
  const size_t n2 = size_t(W2) * H2;
  if (int rc = x.pr_io.ensure(2 * a256(n2) + 2 * a256(n2 * 4) + a256(size_t(W) * H * 4) + 256))
    return rc;
  uint8_t* p = static_cast<uint8_t*>(x.pr_io.ptr);
  uint8_t* L2 = p;
  uint8_t* R2 = L2 + a256(n2);
  float* dh = reinterpret_cast<float*>(R2 + a256(n2));
  float* mh = dh + a256(n2 * 4) / 4;
  float* prior = mh + a256(n2 * 4) / 4;
  int* holes = reinterpret_cast<int*>(reinterpret_cast<uint8_t*>(prior) + a256(size_t(W) * H * 4));
  hipLaunchKernelGGL(k_half, dim3(unsigned((n2 + 255) / 256)), dim3(256), 0, s, dL, dR, W, W2, H2,
                     L2, R2);
  DE_HIP(hipGetLastError());
  if (int rc = x.dense_dev(L2, R2, W2, H2, cc, dh, mh, nullptr, s, 0xF)) return rc;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, holes);
  constexpr int kLines = 256 / kWave;   // lines per workgroup
  hipLaunchKernelGGL(k_fill_rows, dim3((H2 + kLines - 1) / kLines), dim3(256), 0, s, dh, W2, H2);
  hipLaunchKernelGGL(k_fill_cols, dim3((W2 + kLines - 1) / kLines), dim3(256), 0, s, dh, W2, H2,
                     holes);
  DE_HIP(hipGetLastError());
  if (coarse) *coarse = dh;
  int nh = 0;
  DE_HIP(hipMemcpyAsync(&nh, holes, sizeof(int), hipMemcpyDeviceToHost, s));
  DE_HIP(hipStreamSynchronize(s));
  if (used_mask) *used_mask = nh ? 0 : 1;
  if (nh) return x.dense_dev(dL, dR, W, H, c, disp, margin, cand6, s, 0xF);
  hipLaunchKernelGGL(k_upsample, dim3((W * H + 255) / 256), dim3(256), 0, s, dh, W, H, prior);
  DE_HIP(hipGetLastError());
  return prior_dev(x, dL, dR, W, H, c, prior, band, disp, margin, cand6, s);
}

// Stage a host pair in the context's io buffers (through its pinned staging).
static int stage_pair(Ctx& x, const uint8_t* L, const uint8_t* R, int W, int H, bool cand6,
                      uint8_t** dL, uint8_t** dR, float** dd, float** dm, float** dc) {
  const size_t WH = size_t(W) * H;
  if (int rc = x.io_in.ensure(2 * WH)) return rc;
  if (int rc = x.io_out.ensure(WH * 4 * 8)) return rc;
  *dL = static_cast<uint8_t*>(x.io_in.ptr);
  *dR = *dL + WH;
  *dd = static_cast<float*>(x.io_out.ptr);
  *dm = *dd + WH;
  *dc = cand6 ? *dm + WH : nullptr;
  const Xfer in[2] = {{const_cast<uint8_t*>(L), *dL, WH}, {const_cast<uint8_t*>(R), *dR, WH}};
  return x.hio.put(in, 2, x.stream);
}

// Fetch the output planes (and the coarse level, when asked) and wait.
static int fetch_outputs(Ctx& x, int W, int H, float* dd, float* dm, float* dc, float* disp,
                         float* margin, float* cand6, float* coarse = nullptr,
                         const float* dco = nullptr, size_t n_coarse = 0) {
  const size_t WH = size_t(W) * H;
  const Xfer out[4] = {{disp, dd, disp ? WH * 4 : 0},
                       {margin, dm, margin ? WH * 4 : 0},
                       {cand6, dc, cand6 ? WH * 24 : 0},
                       {coarse, const_cast<float*>(dco), coarse && dco ? n_coarse * 4 : 0}};
  return x.hio.get(out, 4, x.stream);
}

}  // namespace de

using namespace de;

extern "C" {

int de_dense_run_prior(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                       const de_cfg* cfg, const float* prior, size_t n_prior, int band,
                       float* disp, float* margin, float* cand6) {
  if (!c || !L || !R || !cfg || !prior || !disp || !margin)
    return fail(DE_ERR_ARG, "de_dense_run_prior: null argument");
  if (W < 1 || H < 1) return fail(DE_ERR_ARG, "image size %dx%d", W, H);
  const size_t WH = size_t(W) * H, half = size_t(W / 2) * (H / 2);
  if (n_prior != WH && (n_prior != half || half == 0))
    return fail(DE_ERR_ARG, "prior has %zu values, expected %zu (full) or %zu (half)", n_prior,
                WH, half);
  for (size_t i = 0; i < n_prior; ++i)
    if (!std::isfinite(prior[i]))
      return fail(DE_ERR_ARG, "prior has holes (value %zu is not finite); fill them first", i);
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  uint8_t *dL, *dR;
  float *dd, *dm, *dc;
  if (int rc = stage_pair(x, L, R, W, H, cand6 != nullptr, &dL, &dR, &dd, &dm, &dc)) return rc;
  if (int rc = x.pr_io.ensure(a256(WH * 4) + a256(half * 4))) return rc;
  float* dprior = static_cast<float*>(x.pr_io.ptr);
  if (n_prior == WH) {
    const Xfer xp = {const_cast<float*>(prior), dprior, WH * 4};
    if (int rc = x.hio.put(&xp, 1, x.stream)) return rc;
  } else {
    float* dhalf = dprior + a256(WH * 4) / 4;
    const Xfer xp = {const_cast<float*>(prior), dhalf, half * 4};
    if (int rc = x.hio.put(&xp, 1, x.stream)) return rc;
    hipLaunchKernelGGL(k_upsample, dim3(unsigned((WH + 255) / 256)), dim3(256), 0, x.stream,
                       dhalf, W, H, dprior);
    DE_HIP(hipGetLastError());
  }
  if (int rc = prior_dev(x, dL, dR, W, H, *cfg, dprior, band, dd, dm, dc, x.stream)) return rc;
  return fetch_outputs(x, W, H, dd, dm, dc, disp, margin, cand6);
}

int de_dense_run_c2f(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                     const de_cfg* cfg, int band, float* disp, float* margin, float* cand6,
                     float* coarse, int* used_mask) {
  if (!c || !L || !R || !cfg || !disp || !margin)
    return fail(DE_ERR_ARG, "de_dense_run_c2f: null argument");
  if (W < 2 || H < 2) return fail(DE_ERR_ARG, "c2f: image %dx%d has no half level", W, H);
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  uint8_t *dL, *dR;
  float *dd, *dm, *dc;
  if (int rc = stage_pair(x, L, R, W, H, cand6 != nullptr, &dL, &dR, &dd, &dm, &dc)) return rc;
  const size_t n2 = size_t(W / 2) * (H / 2);
  const float* dco = nullptr;
  int used = 0;
  if (int rc = c2f_dev(x, dL, dR, W, H, *cfg, band, dd, dm, dc, &dco, &used, x.stream)) return rc;
  if (used_mask) *used_mask = used;
  return fetch_outputs(x, W, H, dd, dm, dc, disp, margin, cand6, coarse, dco, n2);
}

int de_dense_run_c2f_dev(de_ctx* c, const uint8_t* d_L, const uint8_t* d_R, int W, int H,
                         const de_cfg* cfg, int band, float* d_disp, float* d_margin,
                         float* d_cand6, int* used_mask, void* stream) {
  if (!c || !d_L || !d_R || !cfg || !d_disp || !d_margin)
    return fail(DE_ERR_ARG, "de_dense_run_c2f_dev: null argument");
  if (W < 2 || H < 2) return fail(DE_ERR_ARG, "c2f: image %dx%d has no half level", W, H);
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : x.stream;
  int used = 0;
  if (int rc = c2f_dev(x, d_L, d_R, W, H, *cfg, band, d_disp, d_margin, d_cand6, nullptr, &used,
                       s))
    return rc;
  if (used_mask) *used_mask = used;
  return DE_OK;
}

}  // extern "C"
