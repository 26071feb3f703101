// Serving loop: many camera streams of mixed geometry on one GPU.
//
// The reference serves one fixed-geometry stream, one frame in flight
// (`de_dense_cuda --stream`, core/tools/de_dense_cuda.cu:670-770). Here:
//   * a camera stream is opened against a device-memory budget computed from
//     its W x H x D (dense) or W x H x keypoints (keypoint graph); it owns its
//     scratch (cost volume or graph/messages) and `slots` input/output buffer
//     sets, device and pinned host, for its whole life;
//   * submit copies a frame into a free slot (DE_ERR_BUSY when none is free);
//   * one scheduler thread launches continuous batches: whenever a HIP stream
//     is idle it takes the oldest waiting frame of every camera stream whose
//     previous frame has finished (its scratch is then free) and launches them
//     together (one dense batch and one keypoint-graph batch), H2D copies and
//     D2H copies included, on that HIP stream; frames that arrive meanwhile go
//     into the next batch on another HIP stream, so batches overlap;
//   * completions are found by event queries; results stay in the slot's
//     pinned host buffers until released.
// Each frame's result equals a plain call on the same inputs: the batch
// kernels are the single-frame kernels over a batch view.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "de_api_internal.h"

namespace de {
namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

size_t a256(size_t v) { return (v + 255) & ~size_t(255); }

enum SlotState { kFree = 0, kFilling = 1, kPending = 2, kInFlight = 3, kDone = 4 };

struct Slot {
  int state = kFree;
  // pinned host staging and results
  uint8_t* hL = nullptr;
  uint8_t* hR = nullptr;
  int32_t* hkp = nullptr;
  float* hdisp = nullptr;
  int16_t* hq4 = nullptr;
  uint8_t* hconf = nullptr;
  de_knn_out* hout = nullptr;
  // device buffers
  uint8_t* dL = nullptr;
  uint8_t* dR = nullptr;
  int32_t* dkp = nullptr;
  float* ddisp = nullptr;
  float* dmargin = nullptr;
  int16_t* dq4 = nullptr;
  uint8_t* dconf = nullptr;
  de_knn_out* dout = nullptr;
  int n_kp = 0;
  uint64_t tag = 0;
  double t_submit = 0, t_start = 0, t_done = 0;
  int batch = 0, status = DE_OK;
};

struct CamStream {
  int id = -1, kind = DE_STREAM_DENSE, W = 0, H = 0, max_kp = 0;
  de_cfg dcfg;
  de_knn_cfg kcfg;
  void* dev = nullptr;      // one device block: scratch + every slot's buffers
  void* host = nullptr;     // one pinned block: every slot's staging and results
  void* scratch = nullptr;
  size_t bytes = 0;
  bool open = false, busy = false;
  std::vector<Slot> slots;
  std::deque<int> fifo;     // waiting slots, oldest first
};

struct Lane {                // one HIP stream and the batch in flight on it
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<std::pair<int, int>> items;   // (camera stream, slot)
};

}  // namespace
}  // namespace de

using namespace de;

struct de_server {
  de_ctx* ctx = nullptr;
  de_server_cfg cfg;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<CamStream*> cams;
  size_t used = 0;
  std::vector<Lane> lanes;
  std::deque<std::pair<int, int>> done_q;
  std::thread th;
  bool stop = false;
  int launch_error = DE_OK;
  // the dense tables are shared by every dense stream of a server
  bool have_dense = false;
  de_cfg dense_key;
};

namespace de {
namespace {

size_t dense_host_bytes(int W, int H) { return a256(size_t(W) * H) * 2 + a256(size_t(W) * H * 4) + a256(size_t(W) * H * 2) + a256(size_t(W) * H); }
size_t knn_host_bytes(int W, int H, int m) { return a256(size_t(W) * H) * 2 + a256(size_t(m) * 8) + a256(size_t(m) * sizeof(de_knn_out)); }
size_t dense_slot_dev(int W, int H) { return a256(size_t(W) * H) * 2 + a256(size_t(W) * H * 4) * 2 + a256(size_t(W) * H * 2) + a256(size_t(W) * H); }
size_t knn_slot_dev(int W, int H, int m) { return a256(size_t(W) * H) * 2 + a256(size_t(m) * 8) + a256(size_t(m) * sizeof(de_knn_out)); }

// Carve every slot's buffers out of the stream's device and pinned blocks.
void carve(CamStream* c, int nslots) {
  uint8_t* d = static_cast<uint8_t*>(c->dev);
  uint8_t* h = static_cast<uint8_t*>(c->host);
  const size_t WH = size_t(c->W) * c->H;
  c->scratch = d;
  d += a256(c->kind == DE_STREAM_DENSE
                ? scratch_bytes(c->W, c->H, c->dcfg.dmax - c->dcfg.dmin + 1)
                : knn_scratch_bytes(c->W, c->H, c->max_kp));
  auto td = [&](size_t b) { void* r = d; d += a256(b); return r; };
  auto th = [&](size_t b) { void* r = h; h += a256(b); return r; };
  c->slots.assign(size_t(nslots), Slot());
  for (Slot& s : c->slots) {
    s.dL = static_cast<uint8_t*>(td(WH));
    s.dR = static_cast<uint8_t*>(td(WH));
    s.hL = static_cast<uint8_t*>(th(WH));
    s.hR = static_cast<uint8_t*>(th(WH));
    if (c->kind == DE_STREAM_DENSE) {
      s.ddisp = static_cast<float*>(td(WH * 4));
      s.dmargin = static_cast<float*>(td(WH * 4));
      s.dq4 = static_cast<int16_t*>(td(WH * 2));
      s.dconf = static_cast<uint8_t*>(td(WH));
      s.hdisp = static_cast<float*>(th(WH * 4));
      s.hq4 = static_cast<int16_t*>(th(WH * 2));
      s.hconf = static_cast<uint8_t*>(th(WH));
    } else {
      s.dkp = static_cast<int32_t*>(td(size_t(c->max_kp) * 8));
      s.dout = static_cast<de_knn_out*>(td(size_t(c->max_kp) * sizeof(de_knn_out)));
      s.hkp = static_cast<int32_t*>(th(size_t(c->max_kp) * 8));
      s.hout = static_cast<de_knn_out*>(th(size_t(c->max_kp) * sizeof(de_knn_out)));
    }
  }
}

bool same_tables(const de_cfg& a, const de_cfg& b) {
  return a.census_hw == b.census_hw && a.census_hh == b.census_hh && a.sigma_s == b.sigma_s &&
         a.sigma_r == b.sigma_r && a.ad == b.ad && a.ad_trunc == b.ad_trunc;
}

// Launch one batch on lane L (server mutex NOT held; the items' slots are in
// flight, so nothing else touches them).
int launch_batch(de_server* sv, Lane& L, const std::vector<std::pair<CamStream*, int>>& items) {
  Ctx& x = sv->ctx->c;
  std::lock_guard<std::mutex> lk(x.mu);
  DE_HIP(hipSetDevice(x.device));
  std::vector<FrameDesc> fds;
  std::vector<de_knn_frame> kf;
  std::vector<de_knn_cfg> kc;
  std::vector<void*> ks;
  for (const auto& it : items) {
    CamStream* c = it.first;
    Slot& s = c->slots[size_t(it.second)];
    const size_t WH = size_t(c->W) * c->H;
    DE_HIP(hipMemcpyAsync(s.dL, s.hL, WH, hipMemcpyHostToDevice, L.s));
    DE_HIP(hipMemcpyAsync(s.dR, s.hR, WH, hipMemcpyHostToDevice, L.s));
    if (c->kind == DE_STREAM_DENSE) {
      const int D = c->dcfg.dmax - c->dcfg.dmin + 1;
      FrameDesc fd;
      if (int rc = x.frame_desc(c->dcfg, c->W, c->H, s.dL, s.dR, s.ddisp, s.dmargin, nullptr,
                                s.dconf, s.dq4, carve_scratch(c->scratch, c->W, c->H, D), &fd))
        return rc;
      fds.push_back(fd);
    } else {
      if (s.n_kp > 0)
        DE_HIP(hipMemcpyAsync(s.dkp, s.hkp, size_t(s.n_kp) * 8, hipMemcpyHostToDevice, L.s));
      de_knn_frame f;
      f.L = s.dL;
      f.R = s.dR;
      f.W = c->W;
      f.H = c->H;
      f.kp = s.dkp;
      f.n = s.n_kp;
      f.out = s.dout;
      kf.push_back(f);
      kc.push_back(c->kcfg);
      ks.push_back(c->scratch);
    }
  }
  if (!fds.empty())
    if (int rc = x.run_frames(fds.data(), int(fds.size()), L.s, 0xF)) return rc;
  // keypoint frames, one launch sequence per iteration count
  std::vector<char> done(kf.size(), 0);
  for (size_t a = 0; a < kf.size(); ++a) {
    if (done[a]) continue;
    std::vector<de_knn_frame> f1;
    std::vector<de_knn_cfg> c1;
    std::vector<void*> s1;
    for (size_t b = a; b < kf.size(); ++b)
      if (!done[b] && kc[b].iters == kc[a].iters) {
        f1.push_back(kf[b]);
        c1.push_back(kc[b]);
        s1.push_back(ks[b]);
        done[b] = 1;
      }
    if (int rc = knn_launch_frames(x, f1.data(), c1.data(), s1.data(), int(f1.size()), L.s))
      return rc;
  }
  for (const auto& it : items) {
    CamStream* c = it.first;
    Slot& s = c->slots[size_t(it.second)];
    const size_t WH = size_t(c->W) * c->H;
    if (c->kind == DE_STREAM_DENSE) {
      DE_HIP(hipMemcpyAsync(s.hdisp, s.ddisp, WH * 4, hipMemcpyDeviceToHost, L.s));
      DE_HIP(hipMemcpyAsync(s.hq4, s.dq4, WH * 2, hipMemcpyDeviceToHost, L.s));
      DE_HIP(hipMemcpyAsync(s.hconf, s.dconf, WH, hipMemcpyDeviceToHost, L.s));
    } else if (s.n_kp > 0) {
      DE_HIP(hipMemcpyAsync(s.hout, s.dout, size_t(s.n_kp) * sizeof(de_knn_out),
                            hipMemcpyDeviceToHost, L.s));
    }
  }
  DE_HIP(hipEventRecord(L.done, L.s));
  return DE_OK;
}

void scheduler(de_server* sv) {
  (void)hipSetDevice(sv->ctx->c.device);
  std::unique_lock<std::mutex> lk(sv->mu);
  for (;;) {
    // completions
    bool any_inflight = false;
    for (Lane& L : sv->lanes) {
      if (!L.busy) continue;
      const hipError_t q = hipEventQuery(L.done);
      if (q == hipErrorNotReady) {
        any_inflight = true;
        continue;
      }
      const double t = now_s();
      for (const auto& it : L.items) {
        CamStream* c = sv->cams[size_t(it.first)];
        Slot& s = c->slots[size_t(it.second)];
        s.state = kDone;
        s.t_done = t;
        s.status = q == hipSuccess ? DE_OK : DE_ERR_HIP;
        c->busy = false;
        sv->done_q.push_back(it);
      }
      L.items.clear();
      L.busy = false;
      sv->cv_done.notify_all();
    }
    if (sv->stop && !any_inflight) break;
    // launch a batch on an idle lane
    Lane* idle = nullptr;
    for (Lane& L : sv->lanes)
      if (!L.busy) { idle = &L; break; }
    if (idle && !sv->stop) {
      std::vector<std::pair<int, int>> items;
      const int cap = sv->cfg.max_batch > 0 ? sv->cfg.max_batch : 1 << 30;
      for (CamStream* c : sv->cams) {
        if (!c || !c->open || c->busy || c->fifo.empty()) continue;
        if (int(items.size()) >= cap) break;
        const int slot = c->fifo.front();
        c->fifo.pop_front();
        c->busy = true;
        Slot& s = c->slots[size_t(slot)];
        s.state = kInFlight;
        items.emplace_back(c->id, slot);
      }
      if (!items.empty()) {
        const double t = now_s();
        for (const auto& it : items) {
          Slot& s = sv->cams[size_t(it.first)]->slots[size_t(it.second)];
          s.t_start = t;
          s.batch = int(items.size());
        }
        idle->busy = true;
        idle->items = items;
        std::vector<std::pair<CamStream*, int>> work;
        for (const auto& it : items) work.emplace_back(sv->cams[size_t(it.first)], it.second);
        lk.unlock();
        const int rc = launch_batch(sv, *idle, work);
        lk.lock();
        if (rc != DE_OK) {
          // the batch did not launch: report every frame of it as failed
          (void)hipStreamSynchronize(idle->s);
          const double t2 = now_s();
          for (const auto& it : items) {
            CamStream* c = sv->cams[size_t(it.first)];
            Slot& s = c->slots[size_t(it.second)];
            s.state = kDone;
            s.status = rc;
            s.t_done = t2;
            c->busy = false;
            sv->done_q.push_back(it);
          }
          idle->items.clear();
          idle->busy = false;
          sv->launch_error = rc;
          sv->cv_done.notify_all();
        }
        continue;   // look for more work at once
      }
    }
    any_inflight = false;
    for (Lane& L : sv->lanes) any_inflight = any_inflight || L.busy;
    if (any_inflight)
      sv->cv_work.wait_for(lk, std::chrono::microseconds(50));
    else
      sv->cv_work.wait(lk);
  }
}

int open_stream(de_server* sv, CamStream* c, size_t dev_bytes, size_t host_bytes) {
  const int nslots = sv->cfg.slots;
  Ctx& x = sv->ctx->c;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->used + dev_bytes > sv->cfg.mem_budget)
      return fail(DE_ERR_NOMEM, "stream needs %zu bytes of device memory; %zu of the %zu-byte "
                  "budget are in use", dev_bytes, sv->used, sv->cfg.mem_budget);
    sv->used += dev_bytes;
  }
  int rc = DE_OK;
  {
    std::lock_guard<std::mutex> lk(x.mu);
    hipError_t e = hipSetDevice(x.device);
    if (e == hipSuccess) e = hipMalloc(&c->dev, dev_bytes);
    if (e == hipSuccess) e = hipHostMalloc(&c->host, host_bytes, hipHostMallocDefault);
    if (e != hipSuccess) rc = hip_fail(e, "server stream buffers");
    if (rc == DE_OK && c->kind == DE_STREAM_DENSE) {
      const uint16_t *rf, *sc;
      int shift;
      rc = x.tables_for(c->dcfg, &rf, &sc, &shift);
    }
  }
  if (rc != DE_OK) {
    if (c->dev) (void)hipFree(c->dev);
    if (c->host) (void)hipHostFree(c->host);
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->used -= dev_bytes;
    delete c;
    return rc;
  }
  c->bytes = dev_bytes;
  carve(c, nslots);
  std::lock_guard<std::mutex> lk(sv->mu);
  c->id = int(sv->cams.size());
  c->open = true;
  sv->cams.push_back(c);
  return DE_OK;
}

void free_stream(CamStream* c) {
  if (c->dev) (void)hipFree(c->dev);
  if (c->host) (void)hipHostFree(c->host);
  c->dev = c->host = nullptr;
}

}  // namespace
}  // namespace de

extern "C" {

void de_server_default_cfg(de_server_cfg* c) {
  if (!c) return;
  c->mem_budget = size_t(8) << 30;
  c->hip_streams = 4;
  c->slots = 4;
  c->max_batch = 0;
}

int de_server_create(de_ctx* ctx, const de_server_cfg* cfg, de_server** out) {
  if (!ctx || !cfg || !out) return fail(DE_ERR_ARG, "de_server_create: null argument");
  if (cfg->hip_streams < 1 || cfg->hip_streams > 8 || cfg->slots < 1 || cfg->slots > 16 ||
      cfg->max_batch < 0)
    return fail(DE_ERR_ARG, "server: hip_streams 1..8, slots 1..16, max_batch >= 0");
  *out = nullptr;
  de_server* sv = new de_server();
  sv->ctx = ctx;
  sv->cfg = *cfg;
  sv->lanes.resize(size_t(cfg->hip_streams));
  {
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    hipError_t e = hipSetDevice(ctx->c.device);
    for (Lane& L : sv->lanes) {
      if (e == hipSuccess) e = hipStreamCreateWithFlags(&L.s, hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&L.done, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
      for (Lane& L : sv->lanes) {
        if (L.done) (void)hipEventDestroy(L.done);
        if (L.s) (void)hipStreamDestroy(L.s);
      }
      delete sv;
      return hip_fail(e, "server streams");
    }
  }
  sv->th = std::thread(scheduler, sv);
  *out = sv;
  return DE_OK;
}

void de_server_destroy(de_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  if (sv->th.joinable()) sv->th.join();
  std::lock_guard<std::mutex> lk(sv->ctx->c.mu);
  (void)hipSetDevice(sv->ctx->c.device);
  for (Lane& L : sv->lanes) {
    if (L.s) (void)hipStreamSynchronize(L.s);
    if (L.done) (void)hipEventDestroy(L.done);
    if (L.s) (void)hipStreamDestroy(L.s);
  }
  for (CamStream* c : sv->cams) {
    if (!c) continue;
    free_stream(c);
    delete c;
  }
  delete sv;
}

int de_server_open_dense(de_server* sv, int W, int H, const de_cfg* cfg, int* stream_id) {
  if (!sv || !cfg || !stream_id) return fail(DE_ERR_ARG, "open_dense: null argument");
  if (int rc = validate(*cfg, W, H)) return rc;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->have_dense && !same_tables(sv->dense_key, *cfg))
      return fail(DE_ERR_ARG, "dense streams of one server share the census window, sigma_s, "
                  "sigma_r, ad and ad_trunc");
  }
  CamStream* c = new CamStream();
  c->kind = DE_STREAM_DENSE;
  c->W = W;
  c->H = H;
  c->dcfg = *cfg;
  const int D = cfg->dmax - cfg->dmin + 1;
  const size_t dev = a256(scratch_bytes(W, H, D)) + size_t(sv->cfg.slots) * dense_slot_dev(W, H);
  const size_t host = size_t(sv->cfg.slots) * dense_host_bytes(W, H);
  if (int rc = open_stream(sv, c, dev, host)) return rc;
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->have_dense = true;
  sv->dense_key = *cfg;
  *stream_id = c->id;
  return DE_OK;
}

int de_server_open_knn(de_server* sv, int W, int H, int max_kp, const de_knn_cfg* cfg,
                       int* stream_id) {
  if (!sv || !cfg || !stream_id) return fail(DE_ERR_ARG, "open_knn: null argument");
  if (max_kp < 1) return fail(DE_ERR_ARG, "open_knn: max_kp %d", max_kp);
  if (int rc = knn_validate(*cfg, W, H, max_kp)) return rc;
  CamStream* c = new CamStream();
  c->kind = DE_STREAM_KNN;
  c->W = W;
  c->H = H;
  c->max_kp = max_kp;
  c->kcfg = *cfg;
  const size_t dev = a256(knn_scratch_bytes(W, H, max_kp)) +
                     size_t(sv->cfg.slots) * knn_slot_dev(W, H, max_kp);
  const size_t host = size_t(sv->cfg.slots) * knn_host_bytes(W, H, max_kp);
  if (int rc = open_stream(sv, c, dev, host)) return rc;
  *stream_id = c->id;
  return DE_OK;
}

int de_server_close(de_server* sv, int id) {
  if (!sv) return fail(DE_ERR_ARG, "close: null server");
  CamStream* c;
  {
    std::unique_lock<std::mutex> lk(sv->mu);
    if (id < 0 || id >= int(sv->cams.size()) || !sv->cams[size_t(id)] ||
        !sv->cams[size_t(id)]->open)
      return fail(DE_ERR_ARG, "close: no open stream %d", id);
    c = sv->cams[size_t(id)];
    c->open = false;   // nothing new is launched; wait for what is in flight
    c->fifo.clear();
    sv->cv_done.wait(lk, [&] {
      for (const Slot& s : c->slots)
        if (s.state == kInFlight || s.state == kFilling) return false;
      return true;
    });
    // drop its finished results from the queue
    std::deque<std::pair<int, int>> keep;
    for (const auto& it : sv->done_q)
      if (it.first != id) keep.push_back(it);
    sv->done_q.swap(keep);
    sv->used -= c->bytes;
    sv->cams[size_t(id)] = nullptr;
  }
  {
    std::lock_guard<std::mutex> lk(sv->ctx->c.mu);
    (void)hipSetDevice(sv->ctx->c.device);
    free_stream(c);
  }
  delete c;
  return DE_OK;
}

int de_server_memory(de_server* sv, size_t* used, size_t* budget) {
  if (!sv) return fail(DE_ERR_ARG, "memory: null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (used) *used = sv->used;
  if (budget) *budget = sv->cfg.mem_budget;
  return DE_OK;
}

int de_server_submit(de_server* sv, int id, const uint8_t* L, const uint8_t* R,
                     const int32_t* kp, int n_kp, uint64_t tag, int* slot_out) {
  const double t0 = now_s();
  if (!sv || !L || !R) return fail(DE_ERR_ARG, "submit: null argument");
  CamStream* c;
  int slot = -1;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (id < 0 || id >= int(sv->cams.size()) || !sv->cams[size_t(id)] ||
        !sv->cams[size_t(id)]->open)
      return fail(DE_ERR_ARG, "submit: no open stream %d", id);
    c = sv->cams[size_t(id)];
    if (c->kind == DE_STREAM_KNN && (n_kp < 0 || n_kp > c->max_kp || (n_kp > 0 && !kp)))
      return fail(DE_ERR_ARG, "submit: %d keypoints (stream holds %d)", n_kp, c->max_kp);
    for (size_t k = 0; k < c->slots.size(); ++k)
      if (c->slots[k].state == kFree) {
        slot = int(k);
        break;
      }
    if (slot < 0) return fail(DE_ERR_BUSY, "stream %d: every slot is in use", id);
    c->slots[size_t(slot)].state = kFilling;
  }
  Slot& s = c->slots[size_t(slot)];
  const size_t WH = size_t(c->W) * c->H;
  std::memcpy(s.hL, L, WH);
  std::memcpy(s.hR, R, WH);
  if (c->kind == DE_STREAM_KNN) {
    if (n_kp > 0) std::memcpy(s.hkp, kp, size_t(n_kp) * 8);
    s.n_kp = n_kp;
  }
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (!c->open) {   // closed while the frame was being copied in
      s.state = kFree;
      sv->cv_done.notify_all();
      return fail(DE_ERR_ARG, "submit: stream %d was closed", id);
    }
    s.tag = tag;
    s.t_submit = t0;
    s.status = DE_OK;
    s.state = kPending;
    c->fifo.push_back(slot);
  }
  sv->cv_work.notify_one();
  if (slot_out) *slot_out = slot;
  return DE_OK;
}

int de_server_poll(de_server* sv, de_server_result* res, int max, int timeout_ms, int* n_out) {
  if (!sv || !res || !n_out || max < 1) return fail(DE_ERR_ARG, "poll: bad argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  if (sv->done_q.empty() && timeout_ms != 0) {
    auto ready = [&] { return !sv->done_q.empty(); };
    if (timeout_ms < 0) sv->cv_done.wait(lk, ready);
    else sv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
  }
  int n = 0;
  while (n < max && !sv->done_q.empty()) {
    const auto it = sv->done_q.front();
    sv->done_q.pop_front();
    CamStream* c = sv->cams[size_t(it.first)];
    const Slot& s = c->slots[size_t(it.second)];
    de_server_result& r = res[n++];
    std::memset(&r, 0, sizeof(r));
    r.stream_id = it.first;
    r.slot = it.second;
    r.kind = c->kind;
    r.status = s.status;
    r.tag = s.tag;
    r.t_submit = s.t_submit;
    r.t_start = s.t_start;
    r.t_done = s.t_done;
    r.batch_size = s.batch;
    r.W = c->W;
    r.H = c->H;
    r.n = s.n_kp;
    r.disp = s.hdisp;
    r.q4 = s.hq4;
    r.conf = s.hconf;
    r.knn = s.hout;
  }
  *n_out = n;
  return DE_OK;
}

int de_server_release(de_server* sv, int id, int slot) {
  if (!sv) return fail(DE_ERR_ARG, "release: null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (id < 0 || id >= int(sv->cams.size()) || !sv->cams[size_t(id)])
    return fail(DE_ERR_ARG, "release: no stream %d", id);
  CamStream* c = sv->cams[size_t(id)];
  if (slot < 0 || slot >= int(c->slots.size()) || c->slots[size_t(slot)].state != kDone)
    return fail(DE_ERR_ARG, "release: stream %d slot %d holds no result", id, slot);
  c->slots[size_t(slot)].state = kFree;
  return DE_OK;
}

}  // extern "C"
