// Shared host/device definitions for the DoubleEye MI355X stereo path.
//
// Integer representation used on the device. The reference's aggregated score
// is int16 Q14 in [-16384, 16384] (core/tools/de_dense.cpp:398-409). The device
// stores it as the offset bijection v' = v + 16384 in uint16, which is what lets
// one recursive-filter step
//     f = c + ((a*(f_prev - c) + 2^14) >> 15)          (de_dense.cpp:411-430)
// be computed exactly as
//     2*f' = (65536 - 2a)*c' + 2a*f_prev' + 2^15   then   f' = (2*f') >> 16
// i.e. one 24-bit multiply-add off the dependency chain and one 16x16+32
// multiply-add on it, reading f_prev' from the high half of the previous 2*f'.
// Exactness: 2^15*c' is a multiple of 2^15, so floor((2^15 c' + a(f-c) + 2^14)
// / 2^15) = c' + floor((a(f-c) + 2^14) / 2^15); all terms are non-negative and
// below 2^31 + 2^16.
#ifndef DOUBLEEYE_DE_COMMON_H
#define DOUBLEEYE_DE_COMMON_H

#include <stdint.h>

// Explicit address spaces: read-only per-launch data goes through the constant
// space (uniform addresses become scalar loads, freely hoisted), everything a
// kernel writes through the global space (global_* rather than flat_*
// instructions, so LDS waits never wait on outstanding stores).
#if defined(__HIP_DEVICE_COMPILE__)
#define DE_GLOBAL __attribute__((address_space(1)))
#define DE_CONST __attribute__((address_space(4)))
#else
#define DE_GLOBAL
#define DE_CONST
#endif

namespace de {

template <class T>
__host__ __device__ inline DE_GLOBAL T* gptr(T* p) { return (DE_GLOBAL T*)(p); }
template <class T>
__host__ __device__ inline const DE_CONST T* cptr(const T* p) { return (const DE_CONST T*)(p); }

constexpr int kWave = 64;          // CDNA wavefront
constexpr int kQ = 14;
constexpr int kOne = 1 << kQ;      // Q14 unit score
constexpr int kOff = 1 << kQ;      // v' = v + kOff
constexpr int kMaxD = 256;         // 8-bit disparity index in the packed key
constexpr int kMaxG = 4;           // disparity groups of 64 (D <= 256)
constexpr int kMaxW = 4096;        // row solver: 36 bytes of LDS per pixel, 8 pixels per thread
constexpr float kNone = -1e30f;

// Everything one frame needs on the device. All pointers are device memory.
struct FrameDesc {
  const uint8_t* L;      // left  image, W*H
  const uint8_t* R;      // right image, W*H
  uint32_t* lrec;        // left record per pixel, 4 words: census lo, census hi,
                         //   132*L, (2*ax) << 16 | min(65536 - 2*ax, 65535)
  uint64_t* cr;          // right census plane, W*H
  uint32_t* ayt;         // 2*ay, transposed [x][y] (scalar loads down a column)
  uint16_t* vol;         // aggregated volume v', [y][x][Dpad]
  uint16_t* fwd;         // row-pass forward values, per (row, group) ceil(W/8) blocks of
                         //   [64 lanes][8 columns] u16 (1 KB)
  uint32_t* cand;        // per pixel 4 words: key0, key1, (nbm | nbp << 16), 0
  int16_t* rdisp;        // exact reverse match per right pixel (lrc only), W*H
  float* disp;           // outputs, W*H (NaN = no answer)
  float* margin;
  float* cand6;          // nullable: s1, s2, d1, d2, lrc, lrc2 planes
  uint8_t* conf;         // nullable: u8 confidence plane
  int16_t* q4;           // nullable: Q4 disparity plane (-32768 = none)
  const uint16_t* rf;    // 256-entry Q15 coefficient table
  const uint16_t* sc;    // score table of offset values, (nbits+1) << sc_shift
  int W, H, D, Dpad, G, dmin;
  int hw, hh, sc_shift, ad_T;
  int iters, subpixel, fit_eq, lrc;
  int store_vol;         // column pass writes its final values back (tests, exact LRC)
  float lambda, gamma, damping, min_margin, min_ratio;
};

// Mask planes of the prior-masked sweep (--prior / --c2f) of one frame.
struct MaskBufs {
  int* dc;      // [H][W] rounded prior
  int* mx0;     // [H][D] first x of each row's interval (W = none)
  int* mx1;     // [H][D] last x (-1 = none)
  int4* box;    // [D] bounding box of the intervals: ylo, yhi, xlo, xhi (inclusive)
  int4* rect;   // [D] filtered rectangle y0, y1, x0, x1 (half-open); y1 <= y0: skipped
};

// Packed top-2 key: ((v' + 1) << 8) | (255 - k). Integer > orders by value
// descending then k ascending, so the reduction is order independent and
// reproduces the sequential ascending-k strictly-greater insertion
// (core/tools/de_dense.cpp:1286-1293, core/tests/test_top2_merge.cpp).
// 0 means "no candidate".
__host__ __device__ inline uint32_t pack_key(uint32_t vp, int k) {
  return ((vp + 1u) << 8) | uint32_t(255 - k);
}
__host__ __device__ inline int key_k(uint32_t key) { return 255 - int(key & 255u); }
__host__ __device__ inline int key_v(uint32_t key) {
  return int(key >> 8) - 1 - kOff;   // back to signed Q14
}

}  // namespace de

#endif
