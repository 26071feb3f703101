// DEIR -> DEDD stream front end: the packet grammar and pairing of
// `de_dense_cuda --stream` (core/tools/de_dense_cuda.cu:44-59, 670-770), over
// the dense pipeline of this library.
//
//   in:  "DEIR" | w u16 | h u16 | stream u8 | flags u8 | frame u32 | w*h Y8
//   out: "DEDD" | w u16 | h u16 | frame u32 | flags u32 | left Y8 |
//        w*h int16 Q4 disparity (-32768 = no answer) | w*h u8 confidence if flags & 1
//
// Pairing: images wait by frame number (at most 8; beyond that the oldest
// frame numbers are dropped); a second image of the same stream replaces the
// first; stream 1 is the left camera. Fields are little-endian, as the
// reference writes them with memcpy on little-endian hosts.
//
// Pairs are pipelined: up to kInFlight pairs run at once, each in a slot with
// its own device buffers, pinned staging, captured plan and HIP stream, so one
// pair's upload and another's download overlap a third's kernels. Packets
// leave in pairing order. de_stream_feed returns every packet of the pairs
// its bytes completed; de_stream_run_fd keeps pairs in flight while more
// input is already waiting and writes them out before it blocks on a read.
#include <hip/hip_runtime.h>
#include <poll.h>
#include <sys/uio.h>
#include <unistd.h>

#include <cstring>
#include <map>
#include <vector>

#include "de_api_internal.h"

using namespace de;

namespace {

constexpr size_t kHdrIn = 14, kHdrOut = 16;
constexpr uint32_t kDeddConf = 1u;
constexpr size_t kMaxPending = 8;
constexpr int kInFlight = 3;

struct Pending {
  std::vector<uint8_t> img;
  int stream;
};

size_t al256(size_t v) { return (v + 255) & ~size_t(255); }

// Where finished DEDD packets go: the caller's buffer of de_stream_feed (a
// packet that does not fit is dropped and flagged), or a file descriptor the
// fd loop writes to straight from the pinned buffers.
struct Sink {
  uint8_t* p = nullptr;
  size_t cap = 0, len = 0;
  int fd = -1;
  bool overflow = false;
  int put(iovec* iov, int n) {
    size_t bytes = 0;
    for (int i = 0; i < n; ++i) bytes += iov[i].iov_len;
    if (fd < 0) {
      if (bytes > cap - len) {
        overflow = true;
        return DE_OK;
      }
      for (int i = 0; i < n; ++i) {
        std::memcpy(p + len, iov[i].iov_base, iov[i].iov_len);
        len += iov[i].iov_len;
      }
      return DE_OK;
    }
    while (n > 0) {
      const ssize_t w = ::writev(fd, iov, n);
      if (w <= 0) return fail(DE_ERR_ARG, "stream: output closed");
      size_t done = size_t(w);
      while (n > 0 && done >= iov->iov_len) {
        done -= iov->iov_len;
        ++iov;
        --n;
      }
      if (n > 0) {
        iov->iov_base = static_cast<uint8_t*>(iov->iov_base) + done;
        iov->iov_len -= done;
      }
    }
    return DE_OK;
  }
};

// One pair in flight: device L | R | disp | margin | Q4 | conf, pinned
// L | R | Q4 | conf, the pair's captured plan and its stream.
struct Slot {
  void* dev = nullptr;
  void* pin = nullptr;
  de_plan* plan = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  uint32_t frame = 0;
  uint8_t *dL = nullptr, *dR = nullptr, *dconf = nullptr;
  int16_t* dq4 = nullptr;
  uint8_t *hL = nullptr, *hR = nullptr, *hq4 = nullptr, *hconf = nullptr;
};

}  // namespace

struct de_stream {
  de_ctx* ctx = nullptr;
  int W = 0, H = 0;
  de_cfg cfg;
  bool conf = true;
  std::vector<uint8_t> buf;                 // bytes of an incomplete packet
  std::map<uint32_t, Pending> pending;
  long pairs = 0, dropped = 0;
  Slot slot[kInFlight];
  int head = 0;                             // the next slot to launch into
  int busy = 0;                             // slots in flight, the oldest at head - busy
};

namespace {

size_t packet_out_bytes(int W, int H, bool conf) {
  const size_t WH = size_t(W) * H;
  return kHdrOut + WH + WH * 2 + (conf ? WH : 0);
}

// Wait for the oldest pair in flight and append its DEDD packet.
int finish_oldest(de_stream* s, Sink& out) {
  Slot& sl = s->slot[(s->head - s->busy + kInFlight) % kInFlight];
  DE_HIP(hipEventSynchronize(sl.done));
  --s->busy;
  ++s->pairs;
  const size_t WH = size_t(s->W) * s->H;
  uint8_t hdr[kHdrOut];
  const uint16_t pw = uint16_t(s->W), ph = uint16_t(s->H);
  const uint32_t flags = s->conf ? kDeddConf : 0u;
  std::memcpy(hdr, "DEDD", 4);
  std::memcpy(hdr + 4, &pw, 2);
  std::memcpy(hdr + 6, &ph, 2);
  std::memcpy(hdr + 8, &sl.frame, 4);
  std::memcpy(hdr + 12, &flags, 4);
  iovec iov[4] = {{hdr, kHdrOut}, {sl.hL, WH}, {sl.hq4, WH * 2}, {sl.hconf, WH}};
  return out.put(iov, s->conf ? 4 : 3);
}

int finish_all(de_stream* s, Sink& out) {
  while (s->busy > 0)
    if (int rc = finish_oldest(s, out)) return rc;
  return DE_OK;
}

// Launch one pair in the next slot (first finishing the pair that held it).
int launch_pair(de_stream* s, const uint8_t* lft, const uint8_t* rgt, uint32_t frame,
                Sink& out) {
  if (s->busy == kInFlight)
    if (int rc = finish_oldest(s, out)) return rc;
  Slot& sl = s->slot[s->head];
  const size_t WH = size_t(s->W) * s->H;
  DE_HIP(hipSetDevice(s->ctx->c.device));
  std::memcpy(sl.hL, lft, WH);
  DE_HIP(hipMemcpyAsync(sl.dL, sl.hL, WH, hipMemcpyHostToDevice, sl.st));
  std::memcpy(sl.hR, rgt, WH);
  DE_HIP(hipMemcpyAsync(sl.dR, sl.hR, WH, hipMemcpyHostToDevice, sl.st));
  if (int rc = de_plan_launch(sl.plan, sl.st)) return rc;
  DE_HIP(hipMemcpyAsync(sl.hq4, sl.dq4, WH * 2, hipMemcpyDeviceToHost, sl.st));
  if (s->conf) DE_HIP(hipMemcpyAsync(sl.hconf, sl.dconf, WH, hipMemcpyDeviceToHost, sl.st));
  DE_HIP(hipEventRecord(sl.done, sl.st));
  sl.frame = frame;
  s->head = (s->head + 1) % kInFlight;
  ++s->busy;
  return DE_OK;
}

// The --stream pairing rule: an image waits under its frame number until the
// other camera's image of that frame arrives, and then the pair runs (stream
// 1 is the left camera); a newer image from the same camera replaces the
// waiting one; past kMaxPending waiting frames the lowest frame numbers are
// dropped.
int accept_image(de_stream* s, uint32_t num, int cam, const uint8_t* img, Sink& out) {
  const size_t WH = size_t(s->W) * s->H;
  const auto it = s->pending.find(num);
  if (it != s->pending.end() && it->second.stream != cam) {
    const uint8_t* waiting = it->second.img.data();
    const int rc = cam == 1 ? launch_pair(s, img, waiting, num, out)
                            : launch_pair(s, waiting, img, num, out);
    s->pending.erase(it);
    return rc;
  }
  Pending& w = s->pending[num];
  w.img.assign(img, img + WH);
  w.stream = cam;
  for (; s->pending.size() > kMaxPending; ++s->dropped) s->pending.erase(s->pending.begin());
  return DE_OK;
}

// Consume every complete DEIR packet of [p, p + n) (header: "DEIR", w u16,
// h u16, stream u8, flags u8, frame u32; then w*h bytes); returns the bytes
// consumed.
size_t parse(de_stream* s, const uint8_t* p, size_t n, Sink& out, int* rc) {
  const size_t whole = kHdrIn + size_t(s->W) * s->H;
  size_t used = 0;
  *rc = DE_OK;
  while (n - used >= kHdrIn) {
    const uint8_t* h = p + used;
    if (std::memcmp(h, "DEIR", 4) != 0) {
      *rc = fail(DE_ERR_ARG, "stream: lost sync (bad magic)");
      break;
    }
    uint16_t w16, h16;
    uint32_t num;
    std::memcpy(&w16, h + 4, 2);
    std::memcpy(&h16, h + 6, 2);
    std::memcpy(&num, h + 10, 4);
    if (w16 != s->W || h16 != s->H) {
      *rc = fail(DE_ERR_ARG, "stream: packet is %dx%d, the stream was opened for %dx%d",
                 int(w16), int(h16), s->W, s->H);
      break;
    }
    if (n - used < whole) break;   // payload not all here yet
    used += whole;
    if ((*rc = accept_image(s, num, h[8], h + kHdrIn, out)) != DE_OK) break;
  }
  return used;
}

// Feed bytes: parse them in place (after the held bytes of an incomplete
// packet, if any) and hold what is left of an incomplete packet. With
// `finish`, also wait for every pair in flight.
int drain(de_stream* s, const uint8_t* data, size_t n, Sink& out, bool finish) {
  int rc = DE_OK;
  if (s->buf.empty()) {
    const size_t used = parse(s, data, n, out, &rc);
    s->buf.assign(data + used, data + n);
  } else {
    s->buf.insert(s->buf.end(), data, data + n);
    const size_t used = parse(s, s->buf.data(), s->buf.size(), out, &rc);
    s->buf.erase(s->buf.begin(), s->buf.begin() + long(used));
  }
  if (finish) {
    const int rc2 = finish_all(s, out);
    if (rc == DE_OK) rc = rc2;
  }
  return rc;
}

}  // namespace

extern "C" {

size_t de_stream_packet_bytes(int W, int H, int want_conf) {
  return (W > 0 && H > 0) ? packet_out_bytes(W, H, want_conf != 0) : 0;
}

static void release_slots(de_stream* s) {
  (void)hipSetDevice(s->ctx->c.device);
  for (Slot& sl : s->slot) {
    if (sl.st) (void)hipStreamSynchronize(sl.st);
    if (sl.plan) de_plan_destroy(sl.plan);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.st) (void)hipStreamDestroy(sl.st);
    if (sl.dev) (void)hipFree(sl.dev);
    if (sl.pin) (void)hipHostFree(sl.pin);
    sl = Slot();
  }
}

int de_stream_create(de_ctx* ctx, int W, int H, const de_cfg* cfg, int want_conf,
                     de_stream** out) {
  if (!ctx || !cfg || !out) return fail(DE_ERR_ARG, "de_stream_create: null argument");
  if (W < 1 || H < 1 || W > 65535 || H > 65535)
    return fail(DE_ERR_ARG, "stream geometry %dx%d", W, H);
  if (int rc = validate(*cfg, W, H)) return rc;
  *out = nullptr;
  de_stream* s = new de_stream();
  s->ctx = ctx;
  s->W = W;
  s->H = H;
  s->cfg = *cfg;
  s->conf = want_conf != 0;
  const size_t WH = size_t(W) * H;
  const size_t dev_bytes = al256(WH) * 2 + al256(WH * 4) * 2 + al256(WH * 2) + al256(WH);
  const size_t pin_bytes = al256(WH) * 2 + al256(WH * 2) + al256(WH);
  auto bail = [&](int rc) {
    release_slots(s);
    delete s;
    return rc;
  };
  for (Slot& sl : s->slot) {
    hipError_t e = hipSetDevice(ctx->c.device);
    if (e == hipSuccess) e = hipMalloc(&sl.dev, dev_bytes);
    if (e == hipSuccess) e = hipHostMalloc(&sl.pin, pin_bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (e != hipSuccess) return bail(hip_fail(e, "stream slot"));
    uint8_t* p = static_cast<uint8_t*>(sl.dev);
    sl.dL = p;
    sl.dR = p + al256(WH);
    float* ddisp = reinterpret_cast<float*>(p + 2 * al256(WH));
    float* dmargin = reinterpret_cast<float*>(p + 2 * al256(WH) + al256(WH * 4));
    sl.dq4 = reinterpret_cast<int16_t*>(p + 2 * al256(WH) + 2 * al256(WH * 4));
    sl.dconf = p + 2 * al256(WH) + 2 * al256(WH * 4) + al256(WH * 2);
    uint8_t* h = static_cast<uint8_t*>(sl.pin);
    sl.hL = h;
    sl.hR = h + al256(WH);
    sl.hq4 = h + 2 * al256(WH);
    sl.hconf = h + 2 * al256(WH) + al256(WH * 2);
    de_pair pr;
    std::memset(&pr, 0, sizeof(pr));
    pr.L = sl.dL;
    pr.R = sl.dR;
    pr.W = W;
    pr.H = H;
    pr.disp = ddisp;
    pr.margin = dmargin;
    pr.conf = s->conf ? sl.dconf : nullptr;
    pr.q4 = sl.dq4;
    if (int rc = de_plan_create(ctx, &pr, cfg, 1, &sl.plan)) return bail(rc);
  }
  *out = s;
  return DE_OK;
}

void de_stream_destroy(de_stream* s) {
  if (!s) return;
  release_slots(s);
  delete s;
}

int de_stream_feed(de_stream* s, const uint8_t* data, size_t n, uint8_t* out, size_t out_cap,
                   size_t* out_len) {
  if (!s || (n && !data) || !out_len || (out_cap && !out))
    return fail(DE_ERR_ARG, "de_stream_feed: null argument");
  *out_len = 0;
  Sink sk;
  sk.p = out;
  sk.cap = out_cap;
  const int rc = drain(s, data, n, sk, true);
  *out_len = sk.len;
  if (rc == DE_OK && sk.overflow)
    return fail(DE_ERR_ARG, "de_stream_feed: the output capacity %zu holds %zu bytes of packets; "
                "the pairs past it were dropped (size it with de_stream_packet_bytes per pair "
                "fed)", out_cap, sk.len);
  return rc;
}

int de_stream_stats(de_stream* s, long* pairs, long* dropped) {
  if (!s) return fail(DE_ERR_ARG, "de_stream_stats: null stream");
  if (pairs) *pairs = s->pairs;
  if (dropped) *dropped = s->dropped;
  return DE_OK;
}

int de_stream_run_fd(de_ctx* ctx, int in_fd, int out_fd, int W, int H, const de_cfg* cfg,
                     int want_conf, long* pairs, long* dropped) {
  de_stream* s = nullptr;
  if (int rc = de_stream_create(ctx, W, H, cfg, want_conf, &s)) return rc;
  std::vector<uint8_t> in(1 << 20);
  int rc = DE_OK;
  for (;;) {
    const ssize_t got = ::read(in_fd, in.data(), in.size());
    if (got <= 0) break;   // end of input (or a read error): stop like the reference
    // keep pairs in flight only while more input is already waiting
    pollfd pf = {in_fd, POLLIN, 0};
    const bool more = ::poll(&pf, 1, 0) > 0;
    Sink sk;
    sk.fd = out_fd;
    if ((rc = drain(s, in.data(), size_t(got), sk, !more)) != DE_OK) break;
  }
  if (rc == DE_OK) {
    Sink sk;
    sk.fd = out_fd;
    rc = finish_all(s, sk);
  }
  if (pairs) *pairs = s->pairs;
  if (dropped) *dropped = s->dropped;
  de_stream_destroy(s);
  return rc;
}

}  // extern "C"
