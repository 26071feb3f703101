// Batch view shared by every stage kernel: a list of frames and, per stage, the
// prefix of work items each frame contributes. One launch per stage covers all
// frames of the batch; a single pair is a batch of one.
#ifndef DOUBLEEYE_DE_BATCH_H
#define DOUBLEEYE_DE_BATCH_H

#include "de_common.h"

namespace de {

enum Stage { kStagePrep = 0, kStageRow = 1, kStageCol = 2, kStageSolve = 3, kNumStages = 4 };

// Prefetch-depth launch variants of the row and column passes.
constexpr int kNumRowVariants = 4;
constexpr int kNumColVariants = 3;
extern const int kRowVariants[kNumRowVariants];
extern const int kColVariants[kNumColVariants];

// Rows of padding above and below a frame's volume (prefetch runs ahead into it).
constexpr int kVolPadRows = 64;
constexpr int kAytPad = 64;       // words of padding around the transposed coefficients

constexpr int kPrepTileW = 64;   // census/coefficient tile: 64 x 8 pixels
constexpr int kPrepTileH = 8;

// Work items per frame for each stage.
__host__ __device__ inline int stage_items(const FrameDesc& f, int stage) {
  switch (stage) {
    case kStagePrep:
      return ((f.W + kPrepTileW - 1) / kPrepTileW) * ((f.H + kPrepTileH - 1) / kPrepTileH);
    case kStageRow: return f.H;           // one workgroup per row (a wave per 64-disparity group)
    case kStageCol: return f.W;           // one wave per column, all groups
    default: return f.H;                  // one workgroup per row
  }
}

struct BatchView {
  const FrameDesc* frames;   // device array [n]
  const int* start;          // device array [kNumStages][n + 1]
  int n;
};

// Frame owning work item `item` of `stage` (binary search over the prefix).
__device__ inline int find_frame(const BatchView& b, int stage, int item) {
  const int* s = b.start + stage * (b.n + 1);
  int lo = 0, hi = b.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

}  // namespace de

#endif
