// Host-side internals shared by the C-ABI translation units.
#ifndef DOUBLEEYE_DE_API_INTERNAL_H
#define DOUBLEEYE_DE_API_INTERNAL_H

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/doubleeye_hip.h"
#include "de_batch.h"
#include "de_common.h"

namespace de {

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

// Return the library status of a failing HIP call.
#define DE_HIP(call)                                        \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return hip_fail(e_, #call);       \
  } while (0)
int validate(const de_cfg& c, int W, int H);

struct DevBuf {
  void* ptr = nullptr;
  size_t size = 0;
  int ensure(size_t bytes);
  void release();
};

// Pinned host memory (hipHostMalloc), grown on demand.
struct PinBuf {
  void* ptr = nullptr;
  size_t size = 0;
  int ensure(size_t bytes);
  void release();
};

// True when [p, p + n) lies in one allocation made by de_host_alloc: the
// device copies from and to it directly.
bool host_pinned(const void* p, size_t n);

// One plane of a host-pointer entry point: `bytes` between a host and a device
// address.
struct Xfer {
  void* host;
  void* dev;
  size_t bytes;
};

// Host <-> device transfers of the host-pointer entry points. A pageable
// source is copied into pinned staging and its DMA queued at once, so the next
// plane's host copy overlaps it; a pageable destination is fetched in chunks of
// kChunk bytes, each copied out to the caller as soon as its DMA lands. Planes
// in de_host_alloc memory move by DMA alone. get() returns when every plane is
// in host memory.
struct HostIo {
  static constexpr size_t kChunk = size_t(1) << 20;
  PinBuf in, out;
  hipEvent_t in_free = nullptr;       // the last put's DMAs are done with `in`
  std::vector<hipEvent_t> ev;         // one per chunk of a get
  int put(const Xfer* x, int n, hipStream_t s);
  int get(const Xfer* x, int n, hipStream_t s);
  void release();
};

// One frame's device scratch.
struct Scratch {
  uint32_t* lrec;
  uint64_t* cr;
  uint32_t* ayt;
  uint32_t* cand;
  uint16_t* vol;
  uint16_t* fwd;
  int16_t* rdisp;
};
size_t scratch_bytes(int W, int H, int D);
void build_rf_table(float sigma_s, float sigma_r, uint16_t* tbl);
int score_shift(int ad_trunc);
Scratch carve_scratch(void* base, int W, int H, int D);

struct Ctx {
  static constexpr int kSlots = 16;
  struct Slot {
    void* dev = nullptr;
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool used = false;
  };
  int device = -1;
  hipStream_t stream = nullptr;
  std::mutex mu;
  DevBuf scratch, rf_buf, sc_buf, io_in, io_out;
  DevBuf sp_scratch, sp_io;   // sparse matcher: per-frame scratch, staged inputs/outputs
  DevBuf kb_scratch, kb_io;   // keypoint-graph BP: scratch, staged inputs/outputs
  DevBuf ed_scratch, ed_io;   // edge-list MASDA: scratch, staged inputs/outputs
  DevBuf f32_scratch, f32_io; // FP32 full-volume path: scratch, staged inputs/outputs
  DevBuf det_scratch, det_io; // keypoint detector: scratch, staged images and keypoints
  DevBuf pipe_io;             // detect -> match pipeline: descriptors and matches
  DevBuf pr_buf, pr_io;       // prior-masked sweep: mask planes, prior / coarse level
  HostIo hio;                 // pinned staging of the host-pointer entry points
  Slot slots[kSlots];
  int next_slot = 0;
  bool store_vol = false;
  // launch variants (prefetch depths) for the row and column passes; -1 = the
  // tuned or default choice for the shape
  int force_row_variant = -1, force_col_variant = -1;
  int pick_row_variant(int W, int H, int G, int fallback) const;
  int pick_col_variant(int W, int H, int G) const;
  // autotuned variants per shape bucket (W, H, G)
  std::map<std::tuple<int, int, int>, std::pair<int, int>> tuned;
  // optional per-stage timing: hipEvents recorded between the stage launches
  struct EvSet { hipEvent_t e[kNumStages + 1]; };
  bool profiling = false;
  std::vector<EvSet> ev_pending, ev_pool;
  double stage_ms[kNumStages] = {0, 0, 0, 0};
  int stage_runs = 0;
  int take_evset(EvSet* out);
  int drain_events();
  // cached table key
  bool have_tables = false;
  float tk_sigma_s = 0, tk_sigma_r = 0, tk_ad = 0;
  int tk_nbits = 0, tk_T = 0;

  ~Ctx();
  int tables_for(const de_cfg& c, const uint16_t** rf, const uint16_t** sc, int* shift);
  int acquire_slot(size_t bytes, Slot** out);
  int frame_desc(const de_cfg& c, int W, int H, const uint8_t* L, const uint8_t* R,
                 float* disp, float* margin, float* cand6, uint8_t* conf, int16_t* q4,
                 const Scratch& sp, FrameDesc* fd);
  // A batch ready to launch: frames ordered by G, per-stage work prefixes,
  // launch geometry and the chosen variants.
  struct Prepared {
    std::vector<FrameDesc> frames;
    std::vector<int> start;
    size_t lds_row = 0, lds_solve = 0;
    int max_w = 0;                // widest row (the solve launch's pixels per thread)
    int rowv = 0, colv = 0;
    size_t desc_bytes() const { return frames.size() * sizeof(FrameDesc) + start.size() * sizeof(int); }
    void pack(void* dst) const;
  };
  int prepare(const FrameDesc* fds, int n, Prepared* out) const;
  int launch_stages(const Prepared& p, const BatchView& bv, hipStream_t s, int stage_mask,
                    EvSet* ev);
  int run_frames(const FrameDesc* fds, int n, hipStream_t s, int stage_mask);
  int dense_dev(const uint8_t* dL, const uint8_t* dR, int W, int H, const de_cfg& c,
                float* disp, float* margin, float* cand6, hipStream_t s, int stage_mask);
};

void fill_frame_desc(const de_cfg& c, int W, int H, const uint8_t* L, const uint8_t* R,
                     float* disp, float* margin, float* cand6, uint8_t* conf, int16_t* q4,
                     const Scratch& sp, const uint16_t* rf, const uint16_t* sc, bool store_vol,
                     FrameDesc* fd);

// Device address of the allocation-free entry point's scratch array on the
// current device, looked up once per device (de_noalloc.hip).
int noalloc_init_device(int device);

// kernel launchers
size_t row_lds_bytes(int W, int nbits, int sc_shift);
size_t solve_lds_bytes(int W);
hipError_t launch_prep(const BatchView& b, int items, int hw, int hh, hipStream_t s);
hipError_t launch_row(const BatchView& b, int base, int items, int G, int maxW, int T,
                      size_t lds_wave, int variant, hipStream_t s);
size_t row_pass_min_lds(int W, int T, int G);
int device_simds();   // SIMDs of the current device (4 per CU)
hipError_t launch_col(const BatchView& b, int base, int items, int G, bool store, int variant,
                      int min_h, hipStream_t s);
hipError_t launch_solve(const BatchView& b, int items, size_t lds, int maxW, hipStream_t s);
hipError_t launch_lrc_exact(const BatchView& b, int items, hipStream_t s);
hipError_t launch_lrc_plane(const BatchView& b, int items, hipStream_t s);
hipError_t launch_census_plane(const uint8_t* img, int W, int H, int hw, int hh,
                               uint64_t* out, hipStream_t s);

hipError_t launch_masked(const FrameDesc& hf, const FrameDesc* df, const float* prior, int band,
                         const MaskBufs& m, size_t lds_row, bool lrc, hipStream_t s);

// keypoint-graph BP over frames whose scratch blocks (knn_scratch_bytes each)
// the caller provides; launches on s without synchronising
size_t knn_scratch_bytes(int W, int H, int n);
int knn_validate(const de_knn_cfg& c, int W, int H, int n);
int knn_launch_frames(Ctx& x, const de_knn_frame* frames, const de_knn_cfg* cfgs,
                      void* const* scratch, int n, hipStream_t s);

// sparse MASDA of one frame whose keypoints and descriptors are on the device;
// launches on s without synchronising
int sp_validate_cfg(const de_match_cfg& c, int nl, int nr);
int sparse_match_dev(Ctx& x, const de_keypoint* L, const uint64_t* dl, int nl,
                     const de_keypoint* R, const uint64_t* dr, int nr, const de_match_cfg& cfg,
                     de_match* out, int* n_out, de_match_stats* stats, hipStream_t s);

// keypoint detector
int det_validate(const de_detector_cfg& c, int W, int H);
size_t det_max_keypoints(const de_detector_cfg& c, int W, int H);
int det_launch(Ctx& x, const de_detect_frame* frames, const de_detector_cfg* cfgs, int n,
               hipStream_t s, hipEvent_t* ev);

}  // namespace de

// The opaque context handle of the C ABI.
struct de_ctx {
  de::Ctx c;
};

#endif
