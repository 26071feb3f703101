// C-ABI implementation: contexts, resident scratch, host tables and the
// per-stage launches of the dense pipeline. See include/doubleeye_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
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
  fill_frame_desc(c, W, H, L, R, disp, margin, cand6, conf, q4, sp, rf, sc, store_vol, fd);
  return DE_OK;
}

// The descriptor itself, given the frame's device tables (no device calls).
void fill_frame_desc(const de_cfg& c, int W, int H, const uint8_t* L, const uint8_t* R,
                     float* disp, float* margin, float* cand6, uint8_t* conf, int16_t* q4,
                     const Scratch& sp, const uint16_t* rf, const uint16_t* sc, bool store_vol,
                     FrameDesc* fd) {
  std::memset(fd, 0, sizeof(*fd));
  const int D = c.dmax - c.dmin + 1;
  fd->L = L; fd->R = R;
  fd->lrec = sp.lrec; fd->cr = sp.cr; fd->ayt = sp.ayt; fd->vol = sp.vol; fd->cand = sp.cand;
  fd->fwd = sp.fwd;
  fd->disp = disp; fd->margin = margin; fd->cand6 = cand6; fd->conf = conf; fd->q4 = q4;
  fd->rf = rf; fd->sc = sc;
  fd->W = W; fd->H = H; fd->D = D; fd->G = (D + kWave - 1) / kWave; fd->Dpad = fd->G * kWave;
  fd->dmin = c.dmin; fd->hw = c.census_hw; fd->hh = c.census_hh;
  fd->sc_shift = score_shift(c.ad_trunc); fd->ad_T = std::min(255, std::max(1, c.ad_trunc));
  // the exact reverse match reads the final volume, so the column pass stores it
  fd->lrc = (c.lrc && cand6) ? 1 : 0;
  fd->store_vol = (store_vol || fd->lrc) ? 1 : 0;
  fd->rdisp = sp.rdisp;
  fd->iters = c.iters; fd->subpixel = c.subpixel; fd->fit_eq = c.fit_eq;
  fd->lambda = c.lambda; fd->gamma = c.gamma; fd->damping = c.damping;
  fd->min_margin = c.min_margin; fd->min_ratio = c.min_ratio;
}

size_t scratch_bytes(int W, int H, int D) {
  const size_t WH = size_t(W) * H;
  const int Dpad = ((D + kWave - 1) / kWave) * kWave;
  const size_t vol_rows = size_t(H) + 2 * kVolPadRows;
  const size_t fwd = size_t(H) * (Dpad / kWave) * ((W + 7) / 8) * 1024;
  return WH * (16 + 8 + 4 + 16 + 2) + 2 * kAytPad * 4 + vol_rows * W * size_t(Dpad) * 2 + fwd +
         7 * 256;
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
  // ayt carries kAytPad words of padding on each side: the column passes load
  // a chunk's coefficients with one wide load that may run past a column's ends
  s.ayt = reinterpret_cast<uint32_t*>(take(WH * 4 + 2 * kAytPad * 4)) + kAytPad;
  s.cand = reinterpret_cast<uint32_t*>(take(WH * 16));
  s.rdisp = reinterpret_cast<int16_t*>(take(WH * 2));
  // the volume keeps kVolPadRows rows of padding above and below the frame
  const size_t row_elems = size_t(W) * Dpad;
  uint16_t* v = reinterpret_cast<uint16_t*>(take((size_t(H) + 2 * kVolPadRows) * row_elems * 2));
  s.vol = v + kVolPadRows * row_elems;
  // row-pass forward values: per (row, group), ceil(W/8) blocks of 1 KB
  s.fwd = reinterpret_cast<uint16_t*>(take(size_t(H) * (Dpad / kWave) * ((W + 7) / 8) * 1024));
  return s;
}

void Ctx::Prepared::pack(void* dst) const {
  std::memcpy(dst, frames.data(), frames.size() * sizeof(FrameDesc));
  std::memcpy(static_cast<uint8_t*>(dst) + frames.size() * sizeof(FrameDesc), start.data(),
              start.size() * sizeof(int));
}

// Order the frames by disparity-group count (each column-pass bucket, one
// kernel instantiation per G, is then a contiguous range of work items),
// build the per-stage work prefixes and pick the launch variants.
int Ctx::prepare(const FrameDesc* fds_in, int n, Prepared* P) const {
  if (n < 1) return fail(DE_ERR_ARG, "empty batch");
  P->frames.assign(fds_in, fds_in + n);
  std::stable_sort(P->frames.begin(), P->frames.end(),
                   [](const FrameDesc& a, const FrameDesc& b) { return a.G < b.G; });
  const FrameDesc* fds = P->frames.data();
  P->start.assign(size_t(kNumStages) * (n + 1), 0);
  for (int st = 0; st < kNumStages; ++st) {
    int acc = 0;
    for (int i = 0; i < n; ++i) {
      P->start[size_t(st) * (n + 1) + i] = acc;
      acc += stage_items(fds[i], st);
    }
    P->start[size_t(st) * (n + 1) + n] = acc;
  }
  const int hw = fds[0].hw, hh = fds[0].hh;
  int big = 0;
  P->lds_row = P->lds_solve = 0;
  for (int i = 0; i < n; ++i) {
    if (fds[i].hw != hw || fds[i].hh != hh)
      return fail(DE_ERR_ARG, "a batch must share one census window");
    const int nbits = (2 * fds[i].hw + 1) * (2 * fds[i].hh + 1) - 1;
    P->lds_row = std::max(P->lds_row, row_lds_bytes(fds[i].W, nbits, fds[i].sc_shift));
    if (row_pass_min_lds(fds[i].W, fds[i].ad_T, fds[i].G) > 160 * 1024)
      return fail(DE_ERR_ARG, "row of %d pixels too wide for the row pass's LDS", fds[i].W);
    P->lds_solve = std::max(P->lds_solve, solve_lds_bytes(fds[i].W));
    P->max_w = std::max(P->max_w, fds[i].W);
    if (size_t(fds[i].W) * fds[i].H * fds[i].D > size_t(fds[big].W) * fds[big].H * fds[big].D)
      big = i;
  }
  if (P->lds_row > 160 * 1024 || P->lds_solve > 160 * 1024)
    return fail(DE_ERR_ARG, "row too wide for LDS (row %zu, solve %zu bytes)", P->lds_row,
                P->lds_solve);
  // the variants of the largest frame's shape bucket drive the launch
  // untuned row pass: the register-resident kernel, with LDS forward slots
  // when every row wave of the launch fits one per SIMD, without them when the
  // launch takes several rounds (its workgroups then hold little LDS, so a
  // round's waves start beside the previous round's last ones). Measured on
  // the 64-pair batch: 9.77 ms without slots, 9.81 with, 10.29 with the
  // streaming pass (two waves a SIMD).
  size_t row_waves = 0;
  for (int i = 0; i < n; ++i) row_waves += size_t(fds[i].H) * fds[i].G;
  P->rowv = pick_row_variant(fds[big].W, fds[big].H, fds[big].G,
                             row_waves <= size_t(device_simds()) ? 0 : 1);
  P->colv = pick_col_variant(fds[big].W, fds[big].H, fds[big].G);
  return DE_OK;
}

int Ctx::launch_stages(const Prepared& P, const BatchView& bv, hipStream_t s, int stage_mask,
                       EvSet* ev) {
  const int n = int(P.frames.size());
  const FrameDesc* fds = P.frames.data();
  auto items = [&](int st) { return P.start[size_t(st) * (n + 1) + n]; };
  hipError_t e;
  if (ev) DE_HIP(hipEventRecord(ev->e[0], s));
  if ((stage_mask & 1) && (e = launch_prep(bv, items(kStagePrep), fds[0].hw, fds[0].hh, s)) != hipSuccess)
    return hip_fail(e, "k_prep");
  if (ev) DE_HIP(hipEventRecord(ev->e[1], s));
  // row and column passes: one launch per disparity-group count (frames are
  // sorted by G)
  auto per_g = [&](int st, auto&& launch) -> hipError_t {
    for (int i = 0; i < n;) {
      int j = i;
      while (j < n && fds[j].G == fds[i].G) ++j;
      const int base = P.start[size_t(st) * (n + 1) + i];
      const int cnt = P.start[size_t(st) * (n + 1) + j] - base;
      if (hipError_t r = launch(i, j, base, cnt)) return r;
      i = j;
    }
    return hipSuccess;
  };
  if ((stage_mask & 2) &&
      (e = per_g(kStageRow, [&](int i, int j, int base, int cnt) {
         int maxw = 0;
         for (int q = i; q < j; ++q) maxw = std::max(maxw, fds[q].W);
         return launch_row(bv, base, cnt, fds[i].G, maxw, fds[i].ad_T, P.lds_row, P.rowv, s);
       })) != hipSuccess)
    return hip_fail(e, "k_row");
  if (ev) DE_HIP(hipEventRecord(ev->e[2], s));
  if ((stage_mask & 4) &&
      (e = per_g(kStageCol, [&](int i, int j, int base, int cnt) {
         bool store = false;
         int min_h = fds[i].H;
         for (int q = i; q < j; ++q) {
           store = store || fds[q].store_vol != 0;
           min_h = std::min(min_h, fds[q].H);
         }
         return launch_col(bv, base, cnt, fds[i].G, store, P.colv, min_h, s);
       })) != hipSuccess)
    return hip_fail(e, "k_col");
  bool any_lrc = false;
  for (int q = 0; q < n; ++q) any_lrc = any_lrc || fds[q].lrc != 0;
  if ((stage_mask & 4) && any_lrc &&
      (e = launch_lrc_exact(bv, items(kStageSolve), s)) != hipSuccess)
    return hip_fail(e, "k_lrc_exact");
  if (ev) DE_HIP(hipEventRecord(ev->e[3], s));
  if ((stage_mask & 8) && (e = launch_solve(bv, items(kStageSolve), P.lds_solve, P.max_w, s)) != hipSuccess)
    return hip_fail(e, "k_solve");
  if ((stage_mask & 8) && any_lrc &&
      (e = launch_lrc_plane(bv, items(kStageSolve), s)) != hipSuccess)
    return hip_fail(e, "k_lrc_plane");
  if (ev) DE_HIP(hipEventRecord(ev->e[4], s));
  return DE_OK;
}

// Launch every stage over `n` frames (stage_mask selects stages, for tests).
int Ctx::run_frames(const FrameDesc* fds_in, int n, hipStream_t s, int stage_mask) {
  Prepared P;
  if (int rc = prepare(fds_in, n, &P)) return rc;
  const size_t bytes = P.desc_bytes();
  Slot* slot;
  if (int rc = acquire_slot(bytes, &slot)) return rc;
  P.pack(slot->host);
  DE_HIP(hipMemcpyAsync(slot->dev, slot->host, bytes, hipMemcpyHostToDevice, s));
  BatchView bv;
  bv.frames = static_cast<const FrameDesc*>(slot->dev);
  bv.start = reinterpret_cast<const int*>(static_cast<uint8_t*>(slot->dev) + sizeof(FrameDesc) * n);
  bv.n = n;
  const bool prof = profiling && stage_mask == 0xF;
  EvSet ev;
  if (prof) {
    if (int rc = take_evset(&ev)) return rc;
  }
  if (int rc = launch_stages(P, bv, s, stage_mask, prof ? &ev : nullptr)) return rc;
  if (prof) ev_pending.push_back(ev);
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
// column pass; the row pass's fallback is chosen by the launch's size.
int Ctx::pick_row_variant(int W, int H, int G, int fallback) const {
  if (force_row_variant >= 0) return force_row_variant;
  const auto it = tuned.find(std::make_tuple(W, H, G));
  return it != tuned.end() ? it->second.first : fallback;
}
int Ctx::pick_col_variant(int W, int H, int G) const {
  if (force_col_variant >= 0) return force_col_variant;
  const auto it = tuned.find(std::make_tuple(W, H, G));
  return it != tuned.end() ? it->second.second : 1;
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
  for (DevBuf* b : {&scratch, &rf_buf, &sc_buf, &io_in, &io_out, &sp_scratch, &sp_io,
                    &kb_scratch, &kb_io, &ed_scratch, &ed_io, &f32_scratch, &f32_io,
                    &det_scratch, &det_io, &pipe_io, &pr_buf, &pr_io})
    b->release();
  hio.release();
  if (stream) (void)hipStreamDestroy(stream);
}

}  // namespace de

using namespace de;

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
  // set-up of de_dense_run_dev_noalloc's scratch on this device (once per device)
  if (int rc = noalloc_init_device(device)) { delete c; return rc; }
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

int de_alloc(de_ctx* c, size_t bytes, void** d_ptr) {
  if (!c || !d_ptr) return fail(DE_ERR_ARG, "null argument");
  *d_ptr = nullptr;
  std::lock_guard<std::mutex> lk(c->c.mu);
  DE_HIP(hipSetDevice(c->c.device));
  DE_HIP(hipMalloc(d_ptr, std::max(bytes, size_t(1))));
  return DE_OK;
}

int de_free(de_ctx* c, void* d_ptr) {
  if (!c) return fail(DE_ERR_ARG, "null ctx");
  std::lock_guard<std::mutex> lk(c->c.mu);
  DE_HIP(hipSetDevice(c->c.device));
  if (d_ptr) DE_HIP(hipFree(d_ptr));
  return DE_OK;
}

int de_upload(de_ctx* c, void* d_dst, const void* h_src, size_t bytes) {
  if (!c || (!d_dst && bytes) || (!h_src && bytes)) return fail(DE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->c.mu);
  DE_HIP(hipSetDevice(c->c.device));
  DE_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, c->c.stream));
  DE_HIP(hipStreamSynchronize(c->c.stream));
  return DE_OK;
}

int de_download(de_ctx* c, void* h_dst, const void* d_src, size_t bytes) {
  if (!c || (!d_src && bytes) || (!h_dst && bytes)) return fail(DE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->c.mu);
  DE_HIP(hipSetDevice(c->c.device));
  DE_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, c->c.stream));
  DE_HIP(hipStreamSynchronize(c->c.stream));
  return DE_OK;
}

int de_copy_dev(de_ctx* c, void* d_dst, const void* d_src, size_t bytes) {
  if (!c || !d_dst || !d_src) return fail(DE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->c.mu);
  DE_HIP(hipSetDevice(c->c.device));
  DE_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, c->c.stream));
  return DE_OK;
}

// A fill kernel rather than hipMemsetAsync, so the same call is safe inside a
// captured graph.
__global__ void k_fill_bytes(uint8_t* __restrict__ p, uint8_t v, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

int de_fill_dev(de_ctx* c, void* d_dst, int byte, size_t bytes) {
  if (!c || !d_dst) return fail(DE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->c.mu);
  DE_HIP(hipSetDevice(c->c.device));
  const size_t blocks = std::min<size_t>(65536, (bytes + 255) / 256);
  if (blocks)
    hipLaunchKernelGGL(k_fill_bytes, dim3(unsigned(blocks)), dim3(256), 0, c->c.stream,
                       static_cast<uint8_t*>(d_dst), uint8_t(byte), bytes);
  DE_HIP(hipGetLastError());
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

// Host-pointer form: the pair goes up through the context's pinned staging,
// the run is resident, the outputs come back in chunks (HostIo).
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
  const Xfer in[2] = {{const_cast<uint8_t*>(L), dL, WH}, {const_cast<uint8_t*>(R), dR, WH}};
  if (int rc = x.hio.put(in, 2, x.stream)) return rc;
  if (int rc = x.dense_dev(dL, dR, W, H, *cfg, dd, dm, dc, x.stream, stage_mask)) return rc;
  const Xfer out[3] = {{disp, dd, disp ? WH * 4 : 0},
                       {margin, dm, margin ? WH * 4 : 0},
                       {cand6, dc, cand6 ? WH * 24 : 0}};
  return x.hio.get(out, 3, x.stream);
}

int de_dense_run(de_ctx* c, const uint8_t* L, const uint8_t* R, int W, int H,
                 const de_cfg* cfg, float* disp, float* margin, float* cand6) {
  if (!disp || !margin) return fail(DE_ERR_ARG, "de_dense_run: null output");
  if (!c) return fail(DE_ERR_ARG, "null ctx");
  std::lock_guard<std::mutex> lk(c->c.mu);
  return dense_host(c, L, R, W, H, cfg, disp, margin, cand6, 0xF);
}

// ---- batches, plans, autotuner ----

static int batch_check(const de_pair* pairs, const de_cfg* cfgs, int n) {
  if (!pairs || !cfgs || n < 1) return fail(DE_ERR_ARG, "batch: %d pairs", n);
  for (int i = 0; i < n; ++i) {
    const de_pair& p = pairs[i];
    if (!p.L || !p.R || !p.disp || !p.margin)
      return fail(DE_ERR_ARG, "batch pair %d: null input or output", i);
    if (int rc = validate(cfgs[i], p.W, p.H)) return rc;
    const de_cfg& a = cfgs[0];
    const de_cfg& b = cfgs[i];
    if (a.census_hw != b.census_hw || a.census_hh != b.census_hh || a.sigma_s != b.sigma_s ||
        a.sigma_r != b.sigma_r || a.ad != b.ad || a.ad_trunc != b.ad_trunc)
      return fail(DE_ERR_ARG,
                  "batch pair %d: census window, sigma_s, sigma_r, ad and ad_trunc must match "
                  "the batch's first pair", i);
  }
  return DE_OK;
}

static size_t batch_scratch_bytes(const de_pair* pairs, const de_cfg* cfgs, int n,
                                  std::vector<size_t>* offs) {
  size_t total = 0;
  offs->assign(size_t(n), 0);
  for (int i = 0; i < n; ++i) {
    (*offs)[i] = total;
    total += (scratch_bytes(pairs[i].W, pairs[i].H, cfgs[i].dmax - cfgs[i].dmin + 1) + 255) &
             ~size_t(255);
  }
  return total;
}

int de_dense_batch_dev(de_ctx* c, const de_pair* pairs, const de_cfg* cfgs, int n,
                       void* stream) {
  if (!c) return fail(DE_ERR_ARG, "null ctx");
  if (int rc = batch_check(pairs, cfgs, n)) return rc;
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : x.stream;
  std::vector<size_t> offs;
  const size_t total = batch_scratch_bytes(pairs, cfgs, n, &offs);
  if (int rc = x.scratch.ensure(total)) return rc;
  std::vector<FrameDesc> fds(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    const de_pair& p = pairs[i];
    const Scratch sp = carve_scratch(static_cast<uint8_t*>(x.scratch.ptr) + offs[i], p.W, p.H,
                                     cfgs[i].dmax - cfgs[i].dmin + 1);
    if (int rc = x.frame_desc(cfgs[i], p.W, p.H, p.L, p.R, p.disp, p.margin, p.cand6, p.conf,
                              p.q4, sp, &fds[i]))
      return rc;
  }
  return x.run_frames(fds.data(), n, s, 0xF);
}

struct de_plan {
  de_ctx* ctx = nullptr;
  int device = 0;
  DevBuf scratch, desc, tables;
  hipGraph_t graph = nullptr;          // one-graph form: every stage of the batch
  hipGraphExec_t exec = nullptr;
  // Pipelined form: the batch in sub-batches (parts). Part i's prep and row
  // pass (graph 0) run on stream sa in part order; its column pass and solver
  // (graph 1) run on stream sb once its own row pass is done, so part i's
  // column pass overlaps part i + 1's row pass.
  struct Part {
    hipGraph_t g[2] = {nullptr, nullptr};
    hipGraphExec_t e[2] = {nullptr, nullptr};
    hipEvent_t rows_done = nullptr;
  };
  std::vector<Part> parts;
  hipStream_t sa = nullptr, sb = nullptr;
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
};

void de_plan_destroy(de_plan* p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  (void)hipDeviceSynchronize();
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (auto& q : p->parts) {
    for (int h = 0; h < 2; ++h) {
      if (q.e[h]) (void)hipGraphExecDestroy(q.e[h]);
      if (q.g[h]) (void)hipGraphDestroy(q.g[h]);
    }
    if (q.rows_done) (void)hipEventDestroy(q.rows_done);
  }
  if (p->ev_start) (void)hipEventDestroy(p->ev_start);
  if (p->ev_end) (void)hipEventDestroy(p->ev_end);
  if (p->sa) (void)hipStreamDestroy(p->sa);
  if (p->sb) (void)hipStreamDestroy(p->sb);
  p->scratch.release();
  p->desc.release();
  p->tables.release();
  delete p;
}

// Record the stages of `stage_mask` of one prepared batch into a graph.
static int capture_stages(Ctx& x, const Ctx::Prepared& P, const BatchView& bv, int stage_mask,
                          hipGraph_t* graph, hipGraphExec_t* exec) {
  hipStream_t cs = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "plan capture stream");
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  int rc = DE_OK;
  if (e == hipSuccess) {
    rc = x.launch_stages(P, bv, cs, stage_mask, nullptr);
    const hipError_t e2 = hipStreamEndCapture(cs, graph);
    if (rc == DE_OK && e2 != hipSuccess) rc = hip_fail(e2, "hipStreamEndCapture");
  } else {
    rc = hip_fail(e, "hipStreamBeginCapture");
  }
  (void)hipStreamDestroy(cs);
  if (rc != DE_OK) return rc;
  e = hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  return DE_OK;
}

// Sub-batches of a plan. By default none: one graph of every stage. With
// DE_PLAN_PARTS = k in the environment, k parts (at most kMaxPlanParts and n)
// of equal px*disp work (largest frame first into the lightest part), whose
// row and column passes overlap across parts. Measured on the 64-pair mix of
// uneven sizes (census 9x7, 8 iterations): 10.52 ms a batch in one graph,
// 10.54 / 10.74 / 11.32 ms in 2 / 4 / 8 parts -- the batch's row pass alone
// fills the GPU, so overlapping it with another part's column pass only
// shares the same memory paths; one graph stays the default.
constexpr int kMaxPlanParts = 8;

static std::vector<std::vector<int>> plan_parts(const FrameDesc* fds, int n) {
  int np = 1;
  if (const char* env = std::getenv("DE_PLAN_PARTS"))
    if (*env) np = std::atoi(env);
  np = std::max(1, std::min(std::min(np, kMaxPlanParts), n));
  std::vector<std::vector<int>> parts(static_cast<size_t>(np));
  if (np == 1) {
    for (int i = 0; i < n; ++i) parts[0].push_back(i);
    return parts;
  }
  std::vector<int> order(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) order[i] = i;
  auto work = [&](int i) { return double(fds[i].W) * fds[i].H * fds[i].D; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
  std::vector<double> load(static_cast<size_t>(np), 0.0);
  for (int i : order) {
    const int k = int(std::min_element(load.begin(), load.end()) - load.begin());
    parts[k].push_back(i);
    load[k] += work(i);
  }
  return parts;
}

int de_plan_create(de_ctx* c, const de_pair* pairs, const de_cfg* cfgs, int n, de_plan** out) {
  if (!c || !out) return fail(DE_ERR_ARG, "null argument");
  *out = nullptr;
  if (int rc = batch_check(pairs, cfgs, n)) return rc;
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  de_plan* p = new de_plan();
  p->ctx = c;
  p->device = x.device;
  auto bail = [&](int rc) { de_plan_destroy(p); return rc; };
  // private tables: later configuration changes on the context do not touch them
  const de_cfg& c0 = cfgs[0];
  const int nbits = (2 * c0.census_hw + 1) * (2 * c0.census_hh + 1) - 1;
  uint16_t rft[256];
  build_rf_table(c0.sigma_s, c0.sigma_r, rft);
  std::vector<uint16_t> sct;
  build_score_table(nbits, c0.ad, c0.ad_trunc, &sct);
  if (int rc = p->tables.ensure(512 + sct.size() * 2)) return bail(rc);
  hipError_t e = hipMemcpy(p->tables.ptr, rft, 512, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(static_cast<uint8_t*>(p->tables.ptr) + 512, sct.data(), sct.size() * 2,
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) return bail(hip_fail(e, "plan tables"));
  std::vector<size_t> offs;
  const size_t total = batch_scratch_bytes(pairs, cfgs, n, &offs);
  if (int rc = p->scratch.ensure(total)) return bail(rc);
  std::vector<FrameDesc> fds(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    const de_pair& q = pairs[i];
    const Scratch sp = carve_scratch(static_cast<uint8_t*>(p->scratch.ptr) + offs[i], q.W, q.H,
                                     cfgs[i].dmax - cfgs[i].dmin + 1);
    if (int rc = x.frame_desc(cfgs[i], q.W, q.H, q.L, q.R, q.disp, q.margin, q.cand6, q.conf,
                              q.q4, sp, &fds[i]))
      return bail(rc);
    fds[i].rf = static_cast<const uint16_t*>(p->tables.ptr);
    fds[i].sc = reinterpret_cast<const uint16_t*>(static_cast<uint8_t*>(p->tables.ptr) + 512);
  }
  const std::vector<std::vector<int>> groups = plan_parts(fds.data(), n);
  const int np = int(groups.size());
  std::vector<Ctx::Prepared> prep(static_cast<size_t>(np));
  std::vector<size_t> doff(static_cast<size_t>(np));
  size_t dbytes = 0;
  for (int k = 0; k < np; ++k) {
    std::vector<FrameDesc> sub;
    for (int i : groups[k]) sub.push_back(fds[i]);
    if (int rc = x.prepare(sub.data(), int(sub.size()), &prep[k])) return bail(rc);
    doff[k] = dbytes;
    dbytes += (prep[k].desc_bytes() + 255) & ~size_t(255);
  }
  if (int rc = p->desc.ensure(dbytes)) return bail(rc);
  std::vector<uint8_t> host(dbytes);
  for (int k = 0; k < np; ++k) prep[k].pack(host.data() + doff[k]);
  e = hipMemcpy(p->desc.ptr, host.data(), host.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return bail(hip_fail(e, "plan descriptors"));
  auto view = [&](int k) {
    const int m = int(prep[k].frames.size());
    uint8_t* base = static_cast<uint8_t*>(p->desc.ptr) + doff[k];
    BatchView bv;
    bv.frames = reinterpret_cast<const FrameDesc*>(base);
    bv.start = reinterpret_cast<const int*>(base + sizeof(FrameDesc) * m);
    bv.n = m;
    return bv;
  };
  if (np == 1) {
    if (int rc = capture_stages(x, prep[0], view(0), 0xF, &p->graph, &p->exec)) return bail(rc);
    *out = p;
    return DE_OK;
  }
  p->parts.resize(static_cast<size_t>(np));
  for (int k = 0; k < np; ++k) {
    de_plan::Part& q = p->parts[k];
    if (int rc = capture_stages(x, prep[k], view(k), 0x3, &q.g[0], &q.e[0])) return bail(rc);
    if (int rc = capture_stages(x, prep[k], view(k), 0xC, &q.g[1], &q.e[1])) return bail(rc);
    e = hipEventCreateWithFlags(&q.rows_done, hipEventDisableTiming);
    if (e != hipSuccess) return bail(hip_fail(e, "plan event"));
  }
  e = hipStreamCreateWithFlags(&p->sa, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->sb, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_start, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_end, hipEventDisableTiming);
  if (e != hipSuccess) return bail(hip_fail(e, "plan streams"));
  *out = p;
  return DE_OK;
}

int de_plan_parts(const de_plan* p) {
  if (!p) return fail(DE_ERR_ARG, "null plan");
  return p->parts.empty() ? 1 : int(p->parts.size());
}

int de_plan_launch(de_plan* p, void* stream) {
  if (!p || (!p->exec && p->parts.empty())) return fail(DE_ERR_ARG, "null plan");
  DE_HIP(hipSetDevice(p->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : p->ctx->c.stream;
  if (p->parts.empty()) {
    DE_HIP(hipGraphLaunch(p->exec, s));
    return DE_OK;
  }
  // both plan streams start after the work queued on s, and s continues after
  // the last part's solver (every part's column pass waits for its own rows,
  // so the end of stream sb covers stream sa too)
  DE_HIP(hipEventRecord(p->ev_start, s));
  DE_HIP(hipStreamWaitEvent(p->sa, p->ev_start, 0));
  DE_HIP(hipStreamWaitEvent(p->sb, p->ev_start, 0));
  for (auto& q : p->parts) {
    DE_HIP(hipGraphLaunch(q.e[0], p->sa));
    DE_HIP(hipEventRecord(q.rows_done, p->sa));
    DE_HIP(hipStreamWaitEvent(p->sb, q.rows_done, 0));
    DE_HIP(hipGraphLaunch(q.e[1], p->sb));
  }
  DE_HIP(hipEventRecord(p->ev_end, p->sb));
  DE_HIP(hipStreamWaitEvent(s, p->ev_end, 0));
  return DE_OK;
}

int de_autotune(de_ctx* c, int W, int H, const de_cfg* cfg, int reps, int* row_variant,
                int* col_variant) {
  if (!c || !cfg) return fail(DE_ERR_ARG, "null argument");
  if (int rc = validate(*cfg, W, H)) return rc;
  reps = std::max(1, reps);
  // synthetic textured pair (LCG bytes, right = left shifted by D/3)
  const size_t WH = size_t(W) * H;
  std::vector<uint8_t> hl(WH), hr(WH);
  uint32_t st = 12345u;
  for (size_t i = 0; i < WH; ++i) { st = st * 1664525u + 1013904223u; hl[i] = uint8_t(st >> 24); }
  const int sh = std::max(1, (cfg->dmax - cfg->dmin + 1) / 3);
  for (int y = 0; y < H; ++y)
    for (int xx = 0; xx < W; ++xx)
      hr[size_t(y) * W + xx] = hl[size_t(y) * W + std::min(W - 1, xx + sh)];
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  DevBuf io;
  if (int rc = io.ensure(WH * 2 + WH * 8)) return rc;
  uint8_t* dL = static_cast<uint8_t*>(io.ptr);
  uint8_t* dR = dL + WH;
  float* dd = reinterpret_cast<float*>(dR + WH);
  float* dm = dd + WH;
  int rc = DE_OK;
  hipError_t e = hipMemcpy(dL, hl.data(), WH, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dR, hr.data(), WH, hipMemcpyHostToDevice);
  if (e != hipSuccess) { io.release(); return hip_fail(e, "autotune upload"); }
  const bool prof0 = x.profiling;
  const int fr0 = x.force_row_variant, fc0 = x.force_col_variant;
  auto time_stage = [&](int stage, int rv, int cv, double* ms) -> int {
    x.force_row_variant = rv;
    x.force_col_variant = cv;
    if (int r = x.dense_dev(dL, dR, W, H, *cfg, dd, dm, nullptr, x.stream, 0xF)) return r;  // warm
    if (int r = x.drain_events()) return r;
    x.profiling = true;
    for (double& v : x.stage_ms) v = 0.0;
    x.stage_runs = 0;
    for (int i = 0; i < reps; ++i)
      if (int r = x.dense_dev(dL, dR, W, H, *cfg, dd, dm, nullptr, x.stream, 0xF)) return r;
    if (int r = x.drain_events()) return r;
    x.profiling = false;
    *ms = x.stage_ms[stage] / std::max(1, x.stage_runs);
    return DE_OK;
  };
  double best_r = 1e30, best_c = 1e30;
  int br = 2, bc = 1;
  for (int v = 0; v < kNumRowVariants && rc == DE_OK; ++v) {
    double ms = 0;
    rc = time_stage(kStageRow, v, 1, &ms);
    if (rc == DE_OK && ms < best_r) { best_r = ms; br = v; }
  }
  for (int v = 0; v < kNumColVariants && rc == DE_OK; ++v) {
    double ms = 0;
    rc = time_stage(kStageCol, br, v, &ms);
    if (rc == DE_OK && ms < best_c) { best_c = ms; bc = v; }
  }
  x.profiling = prof0;
  for (double& v : x.stage_ms) v = 0.0;
  x.stage_runs = 0;
  x.force_row_variant = fr0;
  x.force_col_variant = fc0;
  (void)hipStreamSynchronize(x.stream);
  io.release();
  if (rc != DE_OK) return rc;
  const int G = (cfg->dmax - cfg->dmin + 1 + kWave - 1) / kWave;
  x.tuned[std::make_tuple(W, H, G)] = std::make_pair(br, bc);
  if (row_variant) *row_variant = br;
  if (col_variant) *col_variant = bc;
  return DE_OK;
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
