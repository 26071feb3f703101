/* CPU oracle for the keypoint-graph belief propagation workload: candidates
 * along the epipolar row by census Hamming distance, a k-nearest-neighbour
 * graph over the keypoints, min-sum loopy BP with a truncated-linear
 * disparity-difference smoothness term, top-2 labels per keypoint.
 * The reference has no such solver; this oracle defines it.
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker, never by the product path. */
#ifndef DOUBLEEYE_ORACLE_KNN_BP_REF_H
#define DOUBLEEYE_ORACLE_KNN_BP_REF_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrcKnnCfg {
  int census_hw, census_hh; /* census window half sizes (9x7: 4, 3) */
  int dmax;                 /* candidate disparities 0..dmax */
  int n_cand;               /* labels kept per keypoint, <= 8 */
  int k_nn;                 /* nearest neighbours per keypoint */
  int iters;                /* BP iterations */
  float w_data;             /* unary cost per Hamming bit */
  float lambda;             /* smoothness cost per pixel of disparity difference */
  int trunc;                /* truncation of the disparity difference, px */
} OrcKnnCfg;

/* Per keypoint: best and second label by (belief, label index). */
typedef struct OrcKnnOut {
  int d1, d2;               /* disparities, -1 when absent */
  float c1, c2;             /* belief costs, +inf when absent */
} OrcKnnOut;

void orc_kb_default_cfg(OrcKnnCfg* c);
/* Labels per keypoint: cand_d/cand_c are [n][8], ncand [n]. Returns 0, or -1
 * if a keypoint's census window leaves the image. */
int orc_kb_candidates(const uint8_t* L, const uint8_t* R, int W, int H, const int32_t* kp,
                      int n, const OrcKnnCfg* cfg, int32_t* cand_d, float* cand_c,
                      int32_t* ncand);
/* Symmetric k-NN graph in CSR form, neighbour lists ascending. off is [n+1];
 * adj holds cap entries. Returns the directed edge count, or -1 if cap is short. */
int orc_kb_knn_graph(const int32_t* kp, int n, int k, int32_t* off, int32_t* adj, int cap);
/* Min-sum BP on a given symmetric graph (sorted CSR lists). beliefs is
 * nullable, [n][8] final beliefs. */
int orc_kb_bp_graph(int n, const int32_t* cand_d, const float* cand_c, const int32_t* ncand,
                    const int32_t* off, const int32_t* adj, const OrcKnnCfg* cfg,
                    OrcKnnOut* out, float* beliefs);
/* The whole workload. */
int orc_kb_run(const uint8_t* L, const uint8_t* R, int W, int H, const int32_t* kp, int n,
               const OrcKnnCfg* cfg, OrcKnnOut* out);
/* Exact min-marginals by enumeration (tiny graphs): mm [n][8], +inf where no
 * label. Energy = sum of unaries + sum over undirected edges of the pairwise
 * term. */
int orc_kb_brute_min_marginals(int n, const int32_t* cand_d, const float* cand_c,
                               const int32_t* ncand, const int32_t* off, const int32_t* adj,
                               const OrcKnnCfg* cfg, double* mm);

#ifdef __cplusplus
}
#endif
#endif
