// de_dense_run_dev_noalloc: the dense pipeline of de_dense_run_dev for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc). It has no counterpart in
// the reference.
//
// Every byte beyond the caller's inputs and outputs lives in one file-scope
// device array, `scratch`, of kNoAllocBytes: the frame's scratch (census and
// coefficient records, the int16 volume, the row pass's forward values,
// candidate records), its two lookup tables and its launch descriptor. The
// array's device address is looked up once per device, when a context is
// created on it (noalloc_init_device), and only that address is ever handed
// to kernels. A set-up kernel writes the descriptor and the tables from its
// own arguments, so a call issues no hipMalloc, no pinned staging and no
// host-to-device copy: kernel launches only.
//
// Every call on a device shares that device's array and it keeps the last
// call's contents; nothing in it has to start at zero (every stage writes
// what the next one reads). Calls that overlap on one device -- two host
// threads, or two streams with no ordering between them -- race on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "de_api_internal.h"
#include "de_batch.h"
#include "de_common.h"

namespace de {
namespace {

// One 848x480 frame at 128 disparities needs about 227 MiB here.
constexpr size_t kNoAllocBytes = size_t(256) << 20;
constexpr size_t kOffRf = 0;            // 256 x u16 filter coefficients
constexpr size_t kOffDesc = 1024;       // FrameDesc, then the stage prefixes
constexpr size_t kOffScore = 4096;      // score table, at most 63 << 8 entries of u16
constexpr size_t kOffFrame = kOffScore + size_t(63 << 8) * 2;
constexpr int kMaxDevices = 64;

static_assert(kOffDesc + sizeof(FrameDesc) + kNumStages * 2 * sizeof(int) <= kOffScore,
              "descriptor region");
static_assert(kOffFrame % 256 == 0, "frame scratch alignment");

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

std::mutex g_mu;
uint8_t* g_ptr[kMaxDevices] = {};

struct SetupArgs {
  FrameDesc fd;
  int start[kNumStages * 2];
  uint16_t rf[256];
  int nbits, shift, T, Tc, wq;
};

// Writes the launch descriptor, the Q15 coefficient table and the graded
// score table (integer closed form of core/tools/de_dense.cpp:1201-1232, the
// same entries the host builds for the other entry points: entry
// (h << shift) + min(|dI|, T) holds v' = v + 16384) into the scratch.
__global__ __launch_bounds__(256) void k_noalloc_setup(SetupArgs a, uint8_t* base) {
  const int t = threadIdx.x;
  if (t == 0) {
    *reinterpret_cast<FrameDesc*>(base + kOffDesc) = a.fd;
    int* st = reinterpret_cast<int*>(base + kOffDesc + sizeof(FrameDesc));
    for (int i = 0; i < kNumStages * 2; ++i) st[i] = a.start[i];
  }
  reinterpret_cast<uint16_t*>(base + kOffRf)[t] = a.rf[t];
  uint16_t* sc = reinterpret_cast<uint16_t*>(base + kOffScore);
  const int half = a.nbits / 2;
  const int n = (a.nbits + 1) << a.shift;
  for (int i = t; i < n; i += 256) {
    const int h = i >> a.shift, v = i & ((1 << a.shift) - 1);
    uint16_t out = uint16_t(kOff);
    if (v <= a.Tc) {
      const int32_t c = ((half - h) * kOne) / half;
      const int32_t ad = kOne - (2 * kOne * min(v, a.T)) / a.T;
      const int32_t s = a.wq ? ((c * (1024 - a.wq) + ad * a.wq) >> 10) : c;
      out = uint16_t(int16_t(s) + kOff);
    }
    sc[i] = out;
  }
}

}  // namespace

int noalloc_init_device(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(DE_ERR_UNSUPPORTED, "device %d: the allocation-free scratch covers devices "
                "0..%d", device, kMaxDevices - 1);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return DE_OK;
  void* p = nullptr;
  // the caller has made `device` current
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  if (!p) return fail(DE_ERR_HIP, "hipGetSymbolAddress(scratch) gave a null address");
  g_ptr[device] = static_cast<uint8_t*>(p);
  return DE_OK;
}

}  // namespace de

using namespace de;

extern "C" {

size_t de_dense_noalloc_limit(void) { return kNoAllocBytes - kOffFrame; }

int de_dense_run_dev_noalloc(de_ctx* c, const uint8_t* dL, const uint8_t* dR, int W, int H,
                             const de_cfg* cfg, float* d_disp, float* d_margin, float* d_cand6,
                             void* stream) {
  if (!c || !dL || !dR || !cfg || !d_disp || !d_margin)
    return fail(DE_ERR_ARG, "de_dense_run_dev_noalloc: null argument");
  if (int rc = validate(*cfg, W, H)) return rc;
  const int D = cfg->dmax - cfg->dmin + 1;
  const size_t need = scratch_bytes(W, H, D);
  if (need > kNoAllocBytes - kOffFrame)
    return fail(DE_ERR_ARG,
                "de_dense_run_dev_noalloc: a %dx%d frame with %d disparities needs %zu bytes of "
                "scratch; the allocation-free entry point holds %zu (use de_dense_run_dev)",
                W, H, D, need, kNoAllocBytes - kOffFrame);
  std::lock_guard<std::mutex> lk(c->c.mu);
  Ctx& x = c->c;
  DE_HIP(hipSetDevice(x.device));
  uint8_t* base = nullptr;
  {
    std::lock_guard<std::mutex> g(g_mu);
    base = (x.device >= 0 && x.device < kMaxDevices) ? g_ptr[x.device] : nullptr;
  }
  if (!base) return fail(DE_ERR_HIP, "no allocation-free scratch on device %d", x.device);
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : x.stream;

  SetupArgs a;
  std::memset(&a, 0, sizeof(a));
  build_rf_table(cfg->sigma_s, cfg->sigma_r, a.rf);
  a.nbits = (2 * cfg->census_hw + 1) * (2 * cfg->census_hh + 1) - 1;
  a.shift = score_shift(cfg->ad_trunc);
  a.T = std::max(1, cfg->ad_trunc);
  a.Tc = std::min(255, a.T);
  a.wq = int32_t(std::max(0.f, std::min(1.f, cfg->ad)) * 1024.f);
  const Scratch sp = carve_scratch(base + kOffFrame, W, H, D);
  fill_frame_desc(*cfg, W, H, dL, dR, d_disp, d_margin, d_cand6, nullptr, nullptr, sp,
                  reinterpret_cast<const uint16_t*>(base + kOffRf),
                  reinterpret_cast<const uint16_t*>(base + kOffScore), false, &a.fd);
  Ctx::Prepared P;
  if (int rc = x.prepare(&a.fd, 1, &P)) return rc;
  for (int i = 0; i < kNumStages * 2; ++i) a.start[i] = P.start[size_t(i)];
  hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(256), 0, s, a, base);
  DE_HIP(hipGetLastError());
  BatchView bv;
  bv.frames = reinterpret_cast<const FrameDesc*>(base + kOffDesc);
  bv.start = reinterpret_cast<const int*>(base + kOffDesc + sizeof(FrameDesc));
  bv.n = 1;
  return x.launch_stages(P, bv, s, 0xF, nullptr);
}

}  // extern "C"
