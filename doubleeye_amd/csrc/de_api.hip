// C-ABI implementation: contexts, resident scratch, host tables and the
// per-stage launches of the dense pipeline. See include/doubleeye_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/doubleeye_hip.h"
#include "de_api_internal.h"
#include "de_batch.h"
#include "de_common.h"

namespace de {

// ---------------------------------------------------------------------------
// errors
static thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(DE_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define DE_HIP(call)                                        \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return hip_fail(e_, #call);       \
  } while (0)

// ---------------------------------------------------------------------------
// device buffers
int DevBuf::ensure(size_t bytes) {
  if (bytes <= size) return DE_OK;
  if (ptr) { (void)hipFree(ptr); ptr = nullptr; size = 0; }
  const size_t want = std::max(bytes, size_t(256));
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) { ptr = nullptr; return hip_fail(e, "hipMalloc"); }
  size = want;
  return DE_OK;
}
void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  size = 0;
}

// ---------------------------------------------------------------------------
// host tables

// Q15 filter coefficient per byte difference, core/tools/de_dense.cpp:531-546
// (same float expression, host libm expf).
void build_rf_table(float sigma_s, float sigma_r, uint16_t* tbl) {
  const float k = -std::sqrt(2.f) / sigma_s;
  const float rs = sigma_s / sigma_r;
  for (int d = 0; d < 256; ++d) {
    const float c = std::exp(k * (1.f + rs * (float(d) / 255.f)));
    tbl[d] = uint16_t(std::min(32767.f, c * 32768.f + 0.5f));
  }
}

int score_shift(int ad_trunc) {
  const int T = std::min(255, std::max(1, ad_trunc));
  int s = 0;
  while ((1 << s) < T + 1) ++s;
  return s;
}

// Graded score table, core/tools/de_dense.cpp:1201-1232: census level
// ((half - h) * ONE) / half (truncating), truncated AD ONE - (2 ONE min(v,T))/T,
// blend (c (1024 - wq) + a wq) >> 10. Entry (h << shift) + min(ad, T) holds the
// offset value v' = v + 16384.
void build_score_table(int nbits, float ad, int ad_trunc, std::vector<uint16_t>* out) {
  const int half = nbits / 2;
  const int T = std::max(1, ad_trunc);
  const int Tc = std::min(255, T);
  const int sh = score_shift(ad_trunc);
  const int32_t wq = int32_t(std::max(0.f, std::min(1.f, ad)) * 1024.f);
  out->assign(size_t(nbits + 1) << sh, uint16_t(kOff));
  for (int h = 0; h <= nbits; ++h) {
    const int32_t c = ((half - h) * kOne) / half;
    for (int v = 0; v <= Tc; ++v) {
      const int32_t a = kOne - (2 * kOne * std::min(v, T)) / T;
      const int32_t s = wq ? ((c * (1024 - wq) + a * wq) >> 10) : c;
      (*out)[(size_t(h) << sh) + v] = uint16_t(int16_t(s) + kOff);
    }
  }
}

int Ctx::tables_for(const de_cfg& c, const uint16_t** rf, const uint16_t** sc, int* shift) {
  const int nbits = (2 * c.census_hw + 1) * (2 * c.census_hh + 1) - 1;
  const bool same = have_tables && tk_sigma_s == c.sigma_s && tk_sigma_r == c.sigma_r &&
                    tk_nbits == nbits && tk_ad == c.ad && tk_T == c.ad_trunc;
  if (!same) {
    uint16_t rft[256];
    build_rf_table(c.sigma_s, c.sigma_r, rft);
    std::vector<uint16_t> sct;
    build_score_table(nbits, c.ad, c.ad_trunc, &sct);
    if (int rc = rf_buf.ensure(sizeof(rft))) return rc;
    if (int rc = sc_buf.ensure(sct.size() * 2)) return rc;
    // the previous tables may still be read by queued kernels
    DE_HIP(hipStreamSynchronize(stream));
    DE_HIP(hipMemcpy(rf_buf.ptr, rft, sizeof(rft), hipMemcpyHostToDevice));
    DE_HIP(hipMemcpy(sc_buf.ptr, sct.data(), sct.size() * 2, hipMemcpyHostToDevice));
    tk_sigma_s = c.sigma_s; tk_sigma_r = c.sigma_r; tk_nbits = nbits;
    tk_ad = c.ad; tk_T = c.ad_trunc; have_tables = true;
  }
  *rf = static_cast<const uint16_t*>(rf_buf.ptr);
  *sc = static_cast<const uint16_t*>(sc_buf.ptr);
  *shift = score_shift(c.ad_trunc);
  return DE_OK;
}

// ---------------------------------------------------------------------------
// descriptor ring: each launch sequence reads its frame descriptors from its own
// device slot, uploaded from pinned staging; a slot is reused only after the
// event recorded behind its last kernel has completed.
int Ctx::acquire_slot(size_t bytes, Slot** out) {
  Slot& s = slots[next_slot];
  next_slot = (next_slot + 1) % kSlots;
  if (s.used) DE_HIP(hipEventSynchronize(s.done));
  if (bytes > s.cap) {
    if (s.dev) (void)hipFree(s.dev);
    if (s.host) (void)hipHostFree(s.host);
    s.dev = nullptr; s.host = nullptr; s.cap = 0;
    const size_t cap = std::max(bytes, size_t(4096));
    DE_HIP(hipMalloc(&s.dev, cap));
    DE_HIP(hipHostMalloc(&s.host, cap, hipHostMallocDefault));
    s.cap = cap;
  }
  if (!s.done) DE_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
  *out = &s;
  return DE_OK;
}

int validate(const de_cfg& c, int W, int H) {
  if (W < 1 || H < 1) return fail(DE_ERR_ARG, "image size %dx%d", W, H);
  if (W > kMaxW) return fail(DE_ERR_ARG, "W=%d exceeds the row solver limit %d", W, kMaxW);
  const int D = c.dmax - c.dmin + 1;
  if (c.dmin < 0 || D < 1 || D > kMaxD)
    return fail(DE_ERR_ARG, "disparity range [%d, %d]: need 0 <= dmin <= dmax, D <= %d",
                c.dmin, c.dmax, kMaxD);
  if (!((c.census_hw == 3 && c.census_hh == 3) || (c.census_hw == 4 && c.census_hh == 3)))
    return fail(DE_ERR_ARG, "census window %dx%d: the dense path supports 7x7 and 9x7",
                2 * c.census_hw + 1, 2 * c.census_hh + 1);
  if (c.iters < 0) return fail(DE_ERR_ARG, "iters %d < 0", c.iters);
  if (!(c.sigma_s > 0.f) || !(c.sigma_r > 0.f))
    return fail(DE_ERR_ARG, "sigma_s and sigma_r must be positive");
  if (c.lrc)
    return fail(DE_ERR_UNSUPPORTED,
                "exact reverse match (--lrc) is not built yet; lrc2 is always written");
  return DE_OK;
}

// Build the descriptor of one frame whose inputs/outputs are device pointers and
// whose scratch lives at the given offsets of the context's scratch buffers.
int Ctx::frame_desc(const de_cfg& c, int W, int H, const uint8_t* L, const uint8_t* R,
                    float* disp, float* margin, float* cand6, uint8_t* conf, int16_t* q4,
                    const Scratch& sp, FrameDesc* fd) {
  const uint16_t *rf, *sc;
  int shift;
  if (int rc = tables_for(c, &rf, &sc, &shift)) return rc;
  std::memset(fd, 0, sizeof(*fd));
  const int D = c.dmax - c.dmin + 1;
  fd->L = L; fd->R = R;
  fd->lrec = sp.lrec; fd->cr = sp.cr; fd->ayt = sp.ayt; fd->vol = sp.vol; fd->cand = sp.cand;
  fd->disp = disp; fd->margin = margin; fd->cand6 = cand6; fd->conf = conf; fd->q4 = q4;
  fd->rf = rf; fd->sc = sc;
  fd->W = W; fd->H = H; fd->D = D; fd->G = (D + kWave - 1) / kWave; fd->Dpad = fd->G * kWave;
  fd->dmin = c.dmin; fd->hw = c.census_hw; fd->hh = c.census_hh;
  fd->sc_shift = shift; fd->ad_T = std::min(255, std::max(1, c.ad_trunc));
  fd->store_vol = store_vol ? 1 : 0;
  fd->iters = c.iters; fd->subpixel = c.subpixel; fd->fit_eq = c.fit_eq; fd->lrc = c.lrc;
  fd->lambda = c.lambda; fd->gamma = c.gamma; fd->damping = c.damping;
  fd->min_margin = c.min_margin; fd->min_ratio = c.min_ratio;
  return DE_OK;
}

size_t scratch_bytes(int W, int H, int D) {
  const size_t WH = size_t(W) * H;
  const int Dpad = ((D + kWave - 1) / kWave) * kWave;
  const size_t vol_rows = size_t(H) + 2 * kVolPadRows;
  return WH * (16 + 8 + 4 + 16) + vol_rows * W * size_t(Dpad) * 2 + 5 * 256;
}

// Carve one frame's scratch out of a device block.
Scratch carve_scratch(void* base, int W, int H, int D) {
  const size_t WH = size_t(W) * H;
  const int Dpad = ((D + kWave - 1) / kWave) * kWave;
  uint8_t* p = static_cast<uint8_t*>(base);
  auto take = [&](size_t bytes) { uint8_t* q = p; p += (bytes + 255) & ~size_t(255); return q; };
  Scratch s;
  s.lrec = reinterpret_cast<uint32_t*>(take(WH * 16));
  s.cr = reinterpret_cast<uint64_t*>(take(WH * 8));
  s.ayt = reinterpret_cast<uint32_t*>(take(WH * 4));
  s.cand = reinterpret_cast<uint32_t*>(take(WH * 16));
  // the volume keeps kVolPadRows rows of padding above and below the frame
  const size_t row_elems = size_t(W) * Dpad;
  uint16_t* v = reinterpret_cast<uint16_t*>(take((size_t(H) + 2 * kVolPadRows) * row_elems * 2));
  s.vol = v + kVolPadRows * row_elems;
  return s;
}

// Launch every stage over `n` frames (stage_mask selects stages, for tests).
int Ctx::run_frames(const FrameDesc* fds_in, int n, hipStream_t s, int stage_mask) {
  if (n < 1) return fail(DE_ERR_ARG, "empty batch");
  // frames ordered by disparity-group count, so each column-pass bucket (one
  // kernel instantiation per G) is a contiguous range of work items
  std::vector<FrameDesc> sorted(fds_in, fds_in + n);
  std::stable_sort(sorted.begin(), sorted.end(),
                   [](const FrameDesc& a, const FrameDesc& b) { return a.G < b.G; });
  const FrameDesc* fds = sorted.data();
  std::vector<int> start(size_t(kNumStages) * (n + 1));
  int gmax = 1;
  size_t lds_row = 0, lds_solve = 0;
  const int hw = fds[0].hw, hh = fds[0].hh;
  for (int st = 0; st < kNumStages; ++st) {
    int acc = 0;
    for (int i = 0; i < n; ++i) {
      start[size_t(st) * (n + 1) + i] = acc;
      acc += stage_items(fds[i], st);
    }
    start[size_t(st) * (n + 1) + n] = acc;
  }
  for (int i = 0; i < n; ++i) {
    if (fds[i].hw != hw || fds[i].hh != hh)
      return fail(DE_ERR_ARG, "a batch must share one census window");
    gmax = std::max(gmax, fds[i].G);
    const int nbits = (2 * fds[i].hw + 1) * (2 * fds[i].hh + 1) - 1;
    lds_row = std::max(lds_row, row_lds_bytes(fds[i].W, nbits, fds[i].sc_shift));
    lds_solve = std::max(lds_solve, solve_lds_bytes(fds[i].W));
  }
  if (lds_row > 160 * 1024 || lds_solve > 160 * 1024)
    return fail(DE_ERR_ARG, "row too wide for LDS (row %zu, solve %zu bytes)", lds_row, lds_solve);
  const size_t bytes = sizeof(FrameDesc) * n + start.size() * sizeof(int);
  Slot* slot;
  if (int rc = acquire_slot(bytes, &slot)) return rc;
  std::memcpy(slot->host, fds, sizeof(FrameDesc) * n);
  std::memcpy(static_cast<uint8_t*>(slot->host) + sizeof(FrameDesc) * n, start.data(),
              start.size() * sizeof(int));
  DE_HIP(hipMemcpyAsync(slot->dev, slot->host, bytes, hipMemcpyHostToDevice, s));
  BatchView bv;
  bv.frames = static_cast<const FrameDesc*>(slot->dev);
  bv.start = reinterpret_cast<const int*>(static_cast<uint8_t*>(slot->dev) + sizeof(FrameDesc) * n);
  bv.n = n;
  auto items = [&](int st) { return start[size_t(st) * (n + 1) + n]; };
  const bool prof = profiling && stage_mask == 0xF;
  EvSet ev;
  if (prof) {
    if (int rc = take_evset(&ev)) return rc;
    DE_HIP(hipEventRecord(ev.e[0], s));
  }
  hipError_t e;
  if ((stage_mask & 1) && (e = launch_prep(bv, items(kStagePrep), hw, hh, s)) != hipSuccess)
    return hip_fail(e, "k_prep");
  if (prof) DE_HIP(hipEventRecord(ev.e[1], s));
  const int rowv = pick_row_variant(fds[0].W, fds[0].H, gmax);
  const int colv = pick_col_variant(fds[0].W, fds[0].H, gmax);
  if ((stage_mask & 2) && (e = launch_row(bv, items(kStageRow), lds_row, rowv, s)) != hipSuccess)
    return hip_fail(e, "k_row");
  if (prof) DE_HIP(hipEventRecord(ev.e[2], s));
  if (stage_mask & 4) {
    for (int i = 0; i < n;) {
      int j = i;
      while (j < n && fds[j].G == fds[i].G) ++j;
      const int base = start[size_t(kStageCol) * (n + 1) + i];
      const int cnt = start[size_t(kStageCol) * (n + 1) + j] - base;
      if ((e = launch_col(bv, base, cnt, fds[i].G, fds[i].store_vol != 0, colv, s)) != hipSuccess)
        return hip_fail(e, "k_col");
      i = j;
    }
  }
  if (prof) DE_HIP(hipEventRecord(ev.e[3], s));
  if ((stage_mask & 8) && (e = launch_solve(bv, items(kStageSolve), lds_solve, s)) != hipSuccess)
    return hip_fail(e, "k_solve");
  if (prof) {
    DE_HIP(hipEventRecord(ev.e[4], s));
    ev_pending.push_back(ev);
  }
  DE_HIP(hipEventRecord(slot->done, s));
  slot->used = true;
  return DE_OK;
}

int Ctx::dense_dev(const uint8_t* dL, const uint8_t* dR, int W, int H, const de_cfg& c,
                   float* disp, float* margin, float* cand6, hipStream_t s, int stage_mask) {
  if (int rc = validate(c, W, H)) return rc;
  const int D = c.dmax - c.dmin + 1;
  if (int rc = scratch.ensure(scratch_bytes(W, H, D))) return rc;
  const Scratch sp = carve_scratch(scratch.ptr, W, H, D);
  FrameDesc fd;
  if (int rc = frame_desc(c, W, H, dL, dR, disp, margin, cand6, nullptr, nullptr, sp, &fd))
    return rc;
  return run_frames(&fd, 1, s, stage_mask);
}

// Defaults until the autotuner has measured a shape: 16 rows in flight in the
// column pass, 32 columns in the row pass's backward scan.
int Ctx::pick_row_variant(int, int, int) const {
  if (force_row_variant >= 0) return force_row_variant;
  return 2;
}
int Ctx::pick_col_variant(int, int, int G) const {
  if (force_col_variant >= 0) return force_col_variant;
  return 1;
}

int Ctx::take_evset(EvSet* out) {
  if (!ev_pool.empty()) { *out = ev_pool.back(); ev_pool.pop_back(); return DE_OK; }
  for (int i = 0; i <= kNumStages; ++i) DE_HIP(hipEventCreate(&out->e[i]));
  return DE_OK;
}

// Fold the elapsed times of every recorded launch sequence into stage_ms.
int Ctx::drain_events() {
  for (EvSet& ev : ev_pending) {
    DE_HIP(hipEventSynchronize(ev.e[kNumStages]));
    for (int st = 0; st < kNumStages; ++st) {
      float ms = 0.f;
      DE_HIP(hipEventElapsedTime(&ms, ev.e[st], ev.e[st + 1]));
      stage_ms[st] += ms;
    }
    ++stage_runs;
    ev_pool.push_back(ev);
  }
  ev_pending.clear();
  return DE_OK;
}

Ctx::~Ctx() {
  if (device >= 0) (void)hipSetDevice(device);
  if (stream) (void)hipStreamSynchronize(stream);
  for (Slot& s : slots) {
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.dev) (void)hipFree(s.dev);
    if (s.host) (void)hipHostFree(s.host);
  }
  for (std::vector<EvSet>* v : {&ev_pending, &ev_pool})
    for (EvSet& ev : *v)
      for (int i = 0; i <= kNumStages; ++i) (void)hipEventDestroy(ev.e[i]);
  for (DevBuf* b : {&scratch, &rf_buf, &sc_buf, &io_in, &io_out})
    b->release();
  if (stream) (void)hipStreamDestroy(stream);
}

}  // namespace de

using namespace de;

struct de_ctx {
  Ctx c;
};

extern "C" {

void de_default_cfg(de_cfg* c) {
  if (!c) return;
  c->dmin = 1; c->dmax = 60; c->iters = 0;
  c->census_hw = 3; c->census_hh = 3;
  c->sigma_s = 8.f; c->sigma_r = 0.2f;
  c->ad = 0.15f; c->ad_trunc = 10;
  c->lambda = -0.1f; c->gamma = -0.1f; c->damping = 0.4f;
  c->min_margin = 0.f; c->min_ratio = 0.f;
  c->subpixel = 1; c->fit_eq = 1; c->lrc = 0;
}

const char* de_last_error(void) { return g_err.c_str(); }

int de_ctx_set_variants(de_ctx* c, int row_variant, int col_variant) {
  if (!c) return fail(DE_ERR_ARG, "null ctx");
  if (row_variant < -1 || row_variant >= kNumRowVariants || col_variant < -1 ||
      col_variant >= kNumColVariants)
    return fail(DE_ERR_ARG, "variant out of range (row %d of %d, col %d of %d)", row_variant,
                kNumRowVariants, col_variant, kNumColVariants);
  std::lock_guard<std::mutex> lk(c->c.mu);
  c->c.force_row_variant = row_variant;
  c->c.force_col_variant = col_variant;
  return DE_OK;
}

int de_ctx_set_profiling(de_ctx* c, int on) {
  if (!c) return fail(DE_ERR_ARG, "null ctx");
  std::lock_guard<std::mutex> lk(c->c.mu);
  DE_HIP(hipSetDevice(c->c.device));
  if (int rc = c->c.drain_events()) return rc;
  c->c.profiling = on != 0;
  for (double& v : c->c.stage_ms) v = 0.0;
  c->c.stage_runs = 0;
  return DE_OK;
}

int de_ctx_stage_ms(de_ctx* c, double* ms4, int* runs) {
  if (!c || !ms4) return fail(DE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->c.mu);
  DE_HIP(hipSetDevice(c->c.device));
  if (int rc = c->c.drain_events()) return rc;
  for (int i = 0; i < kNumStages; ++i) ms4[i] = c->c.stage_ms[i];
  if (runs) *runs = c->c.stage_runs;
  return DE_OK;
}

int de_ctx_create(int device, de_ctx** out) {
  if (!out) return fail(DE_ERR_ARG, "null out");
  *out = nullptr;
  int n = 0;
  DE_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(DE_ERR_ARG, "device %d of %d", device, n);
  DE_HIP(hipSetDevice(device));
  de_ctx* c = new de_ctx();
  c->c.device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipStreamCreate"); }
  *out = c;
  return DE_OK;
}

void de_ctx_destroy(de_ctx* c) { delete c; }

void* de_ctx_stream(de_ctx* c) { return c ? static_cast<void*>(c->c.stream) : nullptr; }

int de_ctx_sync(de_ctx* c) {
  if (!c) return fail(DE_ERR_ARG, "null ctx");
  std::lock_guard<std::mutex> lk(c->c.mu);
  DE_HIP(hipSetDevice(c->c.device));
  DE_HIP(hipStreamSynchronize(c->c.stream));
  return DE_OK;
}

int de_census_dev(de_ctx* c, const uint8_t* d_img, int W, int H, int hw, int hh,
                  uint64_t* d_out, void* stream) {
  if (!c || !d_img || !d_out || W < 1 || H < 1) return fail(DE_ERR_ARG, "de_census_dev args");
  if ((2 * hw + 1) * (2 * hh + 1) - 1 > 64) return fail(DE_ERR_ARG, "census window > 64 bits");
  std::lock_guard<std::mutex> lk(c->c.mu);
  DE_HIP(hipSetDevice(c->c.device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->c.stream;
  const hipError_t e = launch_census_plane(d_img, W, H, hw, hh, d_out, s);
  if (e == hipErrorInvalidValue)
    return fail(DE_ERR_ARG, "census window %dx%d not supported", 2 * hw + 1, 2 * hh + 1);
  if (e != hipSuccess) return hip_fail(e, "k_census_plane");
  return DE_OK;
}

int de_census(de_ctx* c, const uint8_t* img, int W, int H, int hw, int hh, uint64_t* out) {
  if (!c || !img || !out || W < 1 || H < 1) return fail(DE_ERR_ARG, "de_census args");
  const size_t WH = size_t(W) * H;
  {
    std::lock_guard<std::mutex> lk(c->c.mu);
    DE_HIP(hipSetDevice(c->c.device));
    if (int rc = c->c.io_in.ensure(WH)) return rc;
    if (int rc = c->c.io_out.ensure(WH * 8)) return rc;
    DE_HIP(hipMemcpyAsync(c->c.io_in.ptr, img, WH, hipMemcpyHostToDevice, c->c.stream));
  }
  if (int rc = de_census_dev(c, static_cast<uint8_t*>(c->c.io_in.ptr), W, H, hw, hh,
                             static_cast<uint64_t*>(c->c.io_out.ptr), nullptr))
    return rc;
  std::lock_guard<std::mutex> lk(c->c.mu);
  DE_HIP(hipMemcpyAsync(out, c->c.io_out.ptr, WH * 8, hipMemcpyDeviceToHost, c->c.stream));
  DE_HIP(hipStreamSynchronize(c->c.stream));
  return DE_OK;
}

int de_dense_run_dev(de_ctx* c, const uint8_t* dL, const uint8_t* dR, int W, int H,
                     const de_cfg* cfg, float* d_disp, float* d_margin, float* d_cand6,
                     void* stream) {
  if (!c || !dL || !dR || !cfg || !d_disp || !d_margin)
    return fail(DE_ERR_ARG, "de_dense_run_dev: null argument");
  std::lock_guard<std::mutex> lk(c->c.mu);
  DE_HIP(hipSetDevice(c->c.device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->c.stream;
  return c->c.dense_dev(dL, dR, W, H, *cfg, d_disp, d_margin, d_cand6, s, 0xF);
}

// Host-pointer form: stage the pair in the context's io buffers, run, copy back.
static int dense_host(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                      const de_cfg* cfg, float* disp, float* margin, float* cand6,
                      int stage_mask) {
  if (!c || !L || !R || !cfg) return fail(DE_ERR_ARG, "dense: null argument");
  if (int rc = validate(*cfg, W, H)) return rc;
  const size_t WH = size_t(W) * H;
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  if (int rc = x.io_in.ensure(2 * WH)) return rc;
  if (int rc = x.io_out.ensure(WH * 4 * 8)) return rc;
  uint8_t* dL = static_cast<uint8_t*>(x.io_in.ptr);
  uint8_t* dR = dL + WH;
  float* dd = static_cast<float*>(x.io_out.ptr);
  float* dm = dd + WH;
  float* dc = cand6 ? dm + WH : nullptr;
  DE_HIP(hipMemcpyAsync(dL, L, WH, hipMemcpyHostToDevice, x.stream));
  DE_HIP(hipMemcpyAsync(dR, R, WH, hipMemcpyHostToDevice, x.stream));
  if (int rc = x.dense_dev(dL, dR, W, H, *cfg, dd, dm, dc, x.stream, stage_mask)) return rc;
  if (disp) DE_HIP(hipMemcpyAsync(disp, dd, WH * 4, hipMemcpyDeviceToHost, x.stream));
  if (margin) DE_HIP(hipMemcpyAsync(margin, dm, WH * 4, hipMemcpyDeviceToHost, x.stream));
  if (cand6) DE_HIP(hipMemcpyAsync(cand6, dc, WH * 24, hipMemcpyDeviceToHost, x.stream));
  DE_HIP(hipStreamSynchronize(x.stream));
  return DE_OK;
}

int de_dense_run(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                 const de_cfg* cfg, float* disp, float* margin, float* cand6) {
  if (!disp || !margin) return fail(DE_ERR_ARG, "de_dense_run: null output");
  if (!c) return fail(DE_ERR_ARG, "null ctx");
  std::lock_guard<std::mutex> lk(c->c.mu);
  return dense_host(c, L, R, W, H, cfg, disp, margin, cand6, 0xF);
}

int de_dense_volume(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                    const de_cfg* cfg, int k0, int k1, int16_t* out) {
  if (!c || !out || !cfg) return fail(DE_ERR_ARG, "de_dense_volume: null argument");
  const int D = cfg->dmax - cfg->dmin + 1;
  if (k0 < 0 || k1 > D || k0 >= k1) return fail(DE_ERR_ARG, "plane range [%d, %d)", k0, k1);
  std::lock_guard<std::mutex> lk(c->c.mu);
  // full pipeline up to the column pass, with the column pass storing its final
  // (V-bwd) values back into the volume
  c->c.store_vol = true;
  const int rc = dense_host(c, L, R, W, H, cfg, nullptr, nullptr, nullptr, 0x7);
  c->c.store_vol = false;
  if (rc) return rc;
  Ctx& x = c->c;
  const size_t WH = size_t(W) * H;
  const int Dpad = ((D + kWave - 1) / kWave) * kWave;
  const Scratch sp = carve_scratch(x.scratch.ptr, W, H, D);
  std::vector<uint16_t> vol(WH * Dpad);
  DE_HIP(hipMemcpy(vol.data(), sp.vol, vol.size() * 2, hipMemcpyDeviceToHost));
  for (int k = k0; k < k1; ++k) {
    int16_t* P = out + size_t(k - k0) * WH;
    for (size_t i = 0; i < WH; ++i) P[i] = int16_t(int(vol[i * Dpad + k]) - kOff);
  }
  return DE_OK;
}

int de_dense_candidates(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                        const de_cfg* cfg, uint32_t* out) {
  if (!c || !out || !cfg) return fail(DE_ERR_ARG, "de_dense_candidates: null argument");
  std::lock_guard<std::mutex> lk(c->c.mu);
  if (int rc = dense_host(c, L, R, W, H, cfg, nullptr, nullptr, nullptr, 0x7)) return rc;
  Ctx& x = c->c;
  const int D = cfg->dmax - cfg->dmin + 1;
  const Scratch sp = carve_scratch(x.scratch.ptr, W, H, D);
  DE_HIP(hipMemcpy(out, sp.cand, size_t(W) * H * 16, hipMemcpyDeviceToHost));
  return DE_OK;
}

}  // extern "C"
