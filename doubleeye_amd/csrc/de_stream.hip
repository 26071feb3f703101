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
#include <hip/hip_runtime.h>
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

struct Pending {
  std::vector<uint8_t> img;
  int stream;
};

}  // namespace

struct de_stream {
  de_ctx* ctx = nullptr;
  int W = 0, H = 0;
  de_cfg cfg;
  bool conf = true;
  std::vector<uint8_t> buf;                 // bytes not parsed yet
  std::map<uint32_t, Pending> pending;
  long pairs = 0, dropped = 0;
  void* dev = nullptr;
  uint8_t* dL = nullptr;
  uint8_t* dR = nullptr;
  float* ddisp = nullptr;
  float* dmargin = nullptr;
  int16_t* dq4 = nullptr;
  uint8_t* dconf = nullptr;
  std::vector<int16_t> q4;
  std::vector<uint8_t> cf;
};

namespace {

size_t packet_out_bytes(int W, int H, bool conf) {
  const size_t WH = size_t(W) * H;
  return kHdrOut + WH + WH * 2 + (conf ? WH : 0);
}

// Run one pair and append its DEDD packet.
int process_pair(de_stream* s, const std::vector<uint8_t>& lft, const std::vector<uint8_t>& rgt,
                 uint32_t frame, std::vector<uint8_t>* out) {
  const size_t WH = size_t(s->W) * s->H;
  {
    std::lock_guard<std::mutex> lk(s->ctx->c.mu);
    DE_HIP(hipSetDevice(s->ctx->c.device));
    DE_HIP(hipMemcpyAsync(s->dL, lft.data(), WH, hipMemcpyHostToDevice, s->ctx->c.stream));
    DE_HIP(hipMemcpyAsync(s->dR, rgt.data(), WH, hipMemcpyHostToDevice, s->ctx->c.stream));
  }
  de_pair p;
  std::memset(&p, 0, sizeof(p));
  p.L = s->dL;
  p.R = s->dR;
  p.W = s->W;
  p.H = s->H;
  p.disp = s->ddisp;
  p.margin = s->dmargin;
  p.conf = s->conf ? s->dconf : nullptr;
  p.q4 = s->dq4;
  if (int rc = de_dense_batch_dev(s->ctx, &p, &s->cfg, 1, nullptr)) return rc;
  {
    std::lock_guard<std::mutex> lk(s->ctx->c.mu);
    DE_HIP(hipMemcpyAsync(s->q4.data(), s->dq4, WH * 2, hipMemcpyDeviceToHost, s->ctx->c.stream));
    if (s->conf)
      DE_HIP(hipMemcpyAsync(s->cf.data(), s->dconf, WH, hipMemcpyDeviceToHost, s->ctx->c.stream));
    DE_HIP(hipStreamSynchronize(s->ctx->c.stream));
  }
  unsigned char hdr[kHdrOut];
  std::memcpy(hdr, "DEDD", 4);
  const uint16_t pw = uint16_t(s->W), ph = uint16_t(s->H);
  const uint32_t flags = s->conf ? kDeddConf : 0u;
  std::memcpy(hdr + 4, &pw, 2);
  std::memcpy(hdr + 6, &ph, 2);
  std::memcpy(hdr + 8, &frame, 4);
  std::memcpy(hdr + 12, &flags, 4);
  out->insert(out->end(), hdr, hdr + kHdrOut);
  out->insert(out->end(), lft.begin(), lft.end());
  const uint8_t* q = reinterpret_cast<const uint8_t*>(s->q4.data());
  out->insert(out->end(), q, q + WH * 2);
  if (s->conf) out->insert(out->end(), s->cf.begin(), s->cf.end());
  ++s->pairs;
  return DE_OK;
}

// Parse every complete DEIR packet in the buffer; pairs produce DEDD packets.
int drain(de_stream* s, std::vector<uint8_t>* out) {
  const size_t WH = size_t(s->W) * s->H;
  size_t pos = 0;
  int rc = DE_OK;
  while (s->buf.size() - pos >= kHdrIn) {
    const uint8_t* h = s->buf.data() + pos;
    if (std::memcmp(h, "DEIR", 4) != 0) {
      rc = fail(DE_ERR_ARG, "stream: lost sync (bad magic)");
      break;
    }
    uint16_t pw, ph;
    uint32_t num;
    std::memcpy(&pw, h + 4, 2);
    std::memcpy(&ph, h + 6, 2);
    const int st = h[8];
    std::memcpy(&num, h + 10, 4);
    if (int(pw) != s->W || int(ph) != s->H) {
      rc = fail(DE_ERR_ARG, "stream: packet is %dx%d, the stream was opened for %dx%d",
                int(pw), int(ph), s->W, s->H);
      break;
    }
    if (s->buf.size() - pos < kHdrIn + WH) break;   // payload incomplete
    std::vector<uint8_t> img(h + kHdrIn, h + kHdrIn + WH);
    pos += kHdrIn + WH;
    auto it = s->pending.find(num);
    if (it == s->pending.end()) {
      s->pending[num] = Pending{std::move(img), st};
      while (s->pending.size() > kMaxPending) {
        s->pending.erase(s->pending.begin());
        ++s->dropped;
      }
      continue;
    }
    if (it->second.stream == st) {
      it->second.img = std::move(img);
      continue;
    }
    const std::vector<uint8_t>& lft = (st == 1) ? img : it->second.img;
    const std::vector<uint8_t>& rgt = (st == 1) ? it->second.img : img;
    rc = process_pair(s, lft, rgt, num, out);
    s->pending.erase(it);
    if (rc != DE_OK) break;
  }
  s->buf.erase(s->buf.begin(), s->buf.begin() + long(pos));
  return rc;
}

}  // namespace

extern "C" {

size_t de_stream_packet_bytes(int W, int H, int want_conf) {
  return (W > 0 && H > 0) ? packet_out_bytes(W, H, want_conf != 0) : 0;
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
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  const size_t bytes = al(WH) * 2 + al(WH * 4) * 2 + al(WH * 2) + al(WH);
  {
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    hipError_t e = hipSetDevice(ctx->c.device);
    if (e == hipSuccess) e = hipMalloc(&s->dev, bytes);
    if (e != hipSuccess) {
      delete s;
      return hip_fail(e, "stream buffers");
    }
  }
  uint8_t* p = static_cast<uint8_t*>(s->dev);
  s->dL = p;
  p += al(WH);
  s->dR = p;
  p += al(WH);
  s->ddisp = reinterpret_cast<float*>(p);
  p += al(WH * 4);
  s->dmargin = reinterpret_cast<float*>(p);
  p += al(WH * 4);
  s->dq4 = reinterpret_cast<int16_t*>(p);
  p += al(WH * 2);
  s->dconf = p;
  s->q4.resize(WH);
  s->cf.resize(s->conf ? WH : 0);
  *out = s;
  return DE_OK;
}

void de_stream_destroy(de_stream* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->ctx->c.mu);
    (void)hipSetDevice(s->ctx->c.device);
    (void)hipStreamSynchronize(s->ctx->c.stream);
    if (s->dev) (void)hipFree(s->dev);
  }
  delete s;
}

int de_stream_feed(de_stream* s, const uint8_t* data, size_t n, uint8_t* out, size_t out_cap,
                   size_t* out_len) {
  if (!s || (n && !data) || !out_len) return fail(DE_ERR_ARG, "de_stream_feed: null argument");
  *out_len = 0;
  s->buf.insert(s->buf.end(), data, data + n);
  std::vector<uint8_t> o;
  const int rc = drain(s, &o);
  if (o.size() > out_cap) {
    return fail(DE_ERR_ARG, "de_stream_feed: %zu output bytes exceed capacity %zu (size it "
                "with de_stream_packet_bytes per pair fed)", o.size(), out_cap);
  }
  if (!o.empty()) std::memcpy(out, o.data(), o.size());
  *out_len = o.size();
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
    s->buf.insert(s->buf.end(), in.data(), in.data() + got);
    std::vector<uint8_t> o;
    rc = drain(s, &o);
    size_t off = 0;
    while (off < o.size()) {
      const ssize_t w = ::write(out_fd, o.data() + off, o.size() - off);
      if (w <= 0) {
        rc = fail(DE_ERR_ARG, "stream: output closed");
        break;
      }
      off += size_t(w);
    }
    if (rc != DE_OK) break;
  }
  if (pairs) *pairs = s->pairs;
  if (dropped) *dropped = s->dropped;
  de_stream_destroy(s);
  return rc;
}

}  // extern "C"
