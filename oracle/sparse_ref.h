/* CPU oracle for the sparse keypoint MASDA matcher (match_masda) and the
 * build-owned k-NN truncated-linear max-sum BP.
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker, never by the product path. */
#ifndef DOUBLEEYE_ORACLE_SPARSE_REF_H
#define DOUBLEEYE_ORACLE_SPARSE_REF_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* doubleeye::Keypoint (core/include/doubleeye/preproc.hpp:125-130) */
typedef struct OrcKeypoint { float x, y, response, local_std; } OrcKeypoint;

/* doubleeye::Match (core/include/doubleeye/match.hpp:46-65) */
typedef struct OrcMatch {
  int left, right;
  float score, belief, disparity, dy, margin;
} OrcMatch;

/* doubleeye::MatchConfig (match.hpp:136-176) without its two pointers, which
 * are passed separately. */
typedef struct OrcMatchCfg {
  float max_dy, min_disparity, max_disparity;
  int max_candidates;
  float sigma_y, w_y, w_prior, w_smooth;
  int iterations;
  float damping, lambda, gamma, converge_eps;
} OrcMatchCfg;

/* doubleeye::DisparityPrior (match.hpp:81-92) */
typedef struct OrcPrior {
  int cell, cols, rows;
  const float* disparity;
  const float* sigma;
  const uint8_t* valid;
} OrcPrior;

/* doubleeye::SmoothPrior (match.hpp:110-114) */
typedef struct OrcSmooth {
  int n;
  const float* disparity;
  const float* sigma;
  const uint8_t* valid;
} OrcSmooth;

/* doubleeye::MatchStats (match.hpp:178-186) */
typedef struct OrcMatchStats {
  int left, right, candidates, iterations_run, converged;
  float final_max_delta;
  int oscillating;
} OrcMatchStats;

void orc_match_default_cfg(OrcMatchCfg* c);

/* `canonical` = 0 reproduces the reference bit for bit, including libstdc++'s
 * order among exact ties (std::sort / std::partial_sort with the reference's
 * comparators on the same input order). `canonical` = 1 breaks every tie by
 * position instead: right keypoints by (y, index), kept candidates by (score
 * desc, band position), greedy completion by (belief desc, score desc, edge).
 * The two agree whenever no such tie occurs. */
int orc_match_masda(const OrcKeypoint* left, const uint64_t* dl, int nl,
                    const OrcKeypoint* right, const uint64_t* dr, int nr,
                    const OrcMatchCfg* cfg, const OrcPrior* prior,
                    const OrcSmooth* smooth, int canonical,
                    OrcMatch* out, int* n_out, OrcMatchStats* stats);
int orc_match_mutual_nn(const OrcKeypoint* left, const uint64_t* dl, int nl,
                        const OrcKeypoint* right, const uint64_t* dr, int nr,
                        const OrcMatchCfg* cfg, const OrcPrior* prior,
                        const OrcSmooth* smooth, float ratio, int canonical,
                        OrcMatch* out, int* n_out, OrcMatchStats* stats);
/* Candidate edges in the solver's order: for edge e, (ei[e], ej[e], es[e]).
 * Returns the edge count; arrays hold at least nl * max(1, max_candidates). */
int orc_match_candidates(const OrcKeypoint* left, const uint64_t* dl, int nl,
                         const OrcKeypoint* right, const uint64_t* dr, int nr,
                         const OrcMatchCfg* cfg, const OrcPrior* prior,
                         const OrcSmooth* smooth, int canonical,
                         int* ei, int* ej, float* es);
/* Exhaustive maximum-weight matching (factorial; tiny problems only). */
int orc_match_brute_force(const OrcKeypoint* left, const uint64_t* dl, int nl,
                          const OrcKeypoint* right, const uint64_t* dr, int nr,
                          const OrcMatchCfg* cfg, OrcMatch* out, int* n_out);
float orc_matching_objective(const OrcMatch* m, int n, const OrcMatchCfg* cfg,
                             int n_left, int n_right);

/* FAST keypoint detector (core/src/preproc.cpp:371-545) and refine_disparity
 * (core/src/match.cpp:768-810); see detect_ref.cpp. DetectorConfig field order
 * (core/include/doubleeye/preproc.hpp:132-145). */
typedef struct OrcDetectorCfg {
  int grad_window, cell, per_cell;
  float min_response, min_local_std;
  int border, subpixel, fast_threshold, nms_radius;
} OrcDetectorCfg;
void orc_detector_default_cfg(OrcDetectorCfg* c);
int orc_is_fast_corner(const uint8_t* img, int W, int H, int x, int y, int thr);
float orc_shi_tomasi_at(const uint8_t* img, int W, int H, int x, int y, int grad_window);
float orc_local_std_at(const uint8_t* img, int W, int H, int x, int y);
/* counts (nullable): [0] FAST candidates above min_response, [1] survivors of
 * suppression, as DetectProfile reports them (preproc.hpp:191-198). */
int orc_detect_keypoints_fast(const uint8_t* img, int W, int H, const OrcDetectorCfg* cfg,
                              int canonical, OrcKeypoint* out, int cap, int* n_out, int* counts);
void orc_refine_disparity(const uint8_t* L, const uint8_t* R, int W, int H,
                          const OrcKeypoint* kl, int nl, const OrcKeypoint* kr, int nr,
                          OrcMatch* m, int n, int half);

#ifdef __cplusplus
}
#endif
#endif
