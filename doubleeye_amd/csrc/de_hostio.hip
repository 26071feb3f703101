// Host <-> device transfers of the host-pointer entry points: pinned staging
// owned by a context, chunked downloads that hand each chunk to the caller as
// it lands, and de_host_alloc memory that the device reads and writes directly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>

#include "de_api_internal.h"

namespace de {
namespace {

// Allocations made by de_host_alloc: base address -> bytes.
std::mutex g_pin_mu;
std::map<uintptr_t, size_t> g_pinned;

}  // namespace

bool host_pinned(const void* p, size_t n) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  std::lock_guard<std::mutex> lk(g_pin_mu);
  auto it = g_pinned.upper_bound(a);
  if (it == g_pinned.begin()) return false;
  --it;
  return a - it->first <= it->second && n <= it->second - (a - it->first);
}

int PinBuf::ensure(size_t bytes) {
  if (bytes <= size) return DE_OK;
  release();
  const size_t want = std::max(bytes, size_t(1) << 16);
  void* p = nullptr;
  const hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
  if (e != hipSuccess) return hip_fail(e, "hipHostMalloc (staging)");
  ptr = p;
  size = want;
  return DE_OK;
}

void PinBuf::release() {
  if (ptr) (void)hipHostFree(ptr);
  ptr = nullptr;
  size = 0;
}

int HostIo::put(const Xfer* x, int n, hipStream_t s) {
  size_t staged = 0;
  for (int i = 0; i < n; ++i)
    if (x[i].bytes && !host_pinned(x[i].host, x[i].bytes)) staged += (x[i].bytes + 255) & ~size_t(255);
  if (staged) {
    // the previous put's DMAs may still read the staging
    if (in_free) DE_HIP(hipEventSynchronize(in_free));
    if (int rc = in.ensure(staged)) return rc;
  }
  uint8_t* st = static_cast<uint8_t*>(in.ptr);
  for (int i = 0; i < n; ++i) {
    if (!x[i].bytes) continue;
    const void* src = x[i].host;
    if (!host_pinned(x[i].host, x[i].bytes)) {
      std::memcpy(st, x[i].host, x[i].bytes);
      src = st;
      st += (x[i].bytes + 255) & ~size_t(255);
    }
    DE_HIP(hipMemcpyAsync(x[i].dev, src, x[i].bytes, hipMemcpyHostToDevice, s));
  }
  if (staged) {
    if (!in_free) DE_HIP(hipEventCreateWithFlags(&in_free, hipEventDisableTiming));
    DE_HIP(hipEventRecord(in_free, s));
  }
  return DE_OK;
}

int HostIo::get(const Xfer* x, int n, hipStream_t s) {
  size_t staged = 0, chunks = 0;
  for (int i = 0; i < n; ++i)
    if (x[i].bytes && !host_pinned(x[i].host, x[i].bytes)) {
      staged += x[i].bytes;
      chunks += (x[i].bytes + kChunk - 1) / kChunk;
    }
  if (staged)
    if (int rc = out.ensure(staged)) return rc;
  while (ev.size() < chunks) {
    hipEvent_t e = nullptr;
    DE_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ev.push_back(e);
  }
  // queue every DMA first: direct ones, then the staged chunks with an event each
  for (int i = 0; i < n; ++i)
    if (x[i].bytes && host_pinned(x[i].host, x[i].bytes))
      DE_HIP(hipMemcpyAsync(x[i].host, x[i].dev, x[i].bytes, hipMemcpyDeviceToHost, s));
  uint8_t* st = static_cast<uint8_t*>(out.ptr);
  size_t k = 0;
  for (int i = 0; i < n; ++i) {
    if (!x[i].bytes || host_pinned(x[i].host, x[i].bytes)) continue;
    for (size_t o = 0; o < x[i].bytes; o += kChunk) {
      const size_t b = std::min(kChunk, x[i].bytes - o);
      DE_HIP(hipMemcpyAsync(st + o, static_cast<const uint8_t*>(x[i].dev) + o, b,
                            hipMemcpyDeviceToHost, s));
      DE_HIP(hipEventRecord(ev[k++], s));
    }
    st += x[i].bytes;
  }
  // copy each staged chunk out as it lands
  st = static_cast<uint8_t*>(out.ptr);
  k = 0;
  for (int i = 0; i < n; ++i) {
    if (!x[i].bytes || host_pinned(x[i].host, x[i].bytes)) continue;
    for (size_t o = 0; o < x[i].bytes; o += kChunk) {
      const size_t b = std::min(kChunk, x[i].bytes - o);
      DE_HIP(hipEventSynchronize(ev[k++]));
      std::memcpy(static_cast<uint8_t*>(x[i].host) + o, st + o, b);
    }
    st += x[i].bytes;
  }
  DE_HIP(hipStreamSynchronize(s));
  return DE_OK;
}

void HostIo::release() {
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  ev.clear();
  if (in_free) (void)hipEventDestroy(in_free);
  in_free = nullptr;
  in.release();
  out.release();
}

}  // namespace de

using namespace de;

extern "C" {

int de_host_alloc(size_t bytes, void** out) {
  if (!out) return fail(DE_ERR_ARG, "de_host_alloc: null output");
  *out = nullptr;
  if (!bytes) return fail(DE_ERR_ARG, "de_host_alloc: zero bytes");
  void* p = nullptr;
  const hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) return hip_fail(e, "hipHostMalloc");
  {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_pinned[reinterpret_cast<uintptr_t>(p)] = bytes;
  }
  *out = p;
  return DE_OK;
}

int de_host_free(void* p) {
  if (!p) return DE_OK;
  {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    auto it = g_pinned.find(reinterpret_cast<uintptr_t>(p));
    if (it == g_pinned.end())
      return fail(DE_ERR_ARG, "de_host_free: %p was not returned by de_host_alloc", p);
    g_pinned.erase(it);
  }
  DE_HIP(hipHostFree(p));
  return DE_OK;
}

}  // extern "C"
