/* CPU oracle for the DoubleEye dense/sparse MASDA hot path.
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker, never by the product path. */
#ifndef DOUBLEEYE_ORACLE_CPU_REF_H
#define DOUBLEEYE_ORACLE_CPU_REF_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors the default-path fields of doubleeye::Cfg
 * (core/include/doubleeye/dense_solve.hpp:19-97) plus the census window. */
typedef struct OrcCfg {
  int dmin, dmax, iters;
  int census_hw, census_hh;
  float sigma_s, sigma_r;
  float ad;
  int ad_trunc;
  float lambda, gamma, damping, min_margin, min_ratio;
  int subpixel, fit_eq, lrc;
} OrcCfg;

void orc_default_cfg(OrcCfg* c);
int orc_census(const uint8_t* img, int W, int H, int hw, int hh, uint64_t* out);
void orc_rf_table(float sigma_s, float sigma_r, uint16_t* tbl);
void orc_rf_coeffs(const uint8_t* L, int W, int H, float sigma_s, float sigma_r,
                   uint16_t* ax, uint16_t* ay);
void orc_score_tables(int half, float ad, int ad_trunc, int16_t* tbl65,
                      int16_t* adt256, int32_t* wq);
int orc_filtered_planes_i16(const uint8_t* L, const uint8_t* R, int W, int H,
                            const OrcCfg* cfg, int k0, int k1, int16_t* out);
int orc_dense_run(const uint8_t* L, const uint8_t* R, int W, int H,
                  const OrcCfg* cfg, float* disp, float* margin, float* cand6,
                  uint8_t* tie_mask);
int orc_dense_run_full(const uint8_t* L, const uint8_t* R, int W, int H, const OrcCfg* cfg,
                       float* disp, float* margin, uint8_t* tie_mask);
int orc_volume_f32(const uint8_t* L, const uint8_t* R, int W, int H,
                   const OrcCfg* cfg, float* vol);
float orc_confidence(float s1, float s2, float lambda, float lrc, float lrc_tol);
uint8_t orc_confidence_u8(float s1, float s2, float lambda, float lrc, float lrc_tol);
void orc_dedd_planes(const float* disp, const float* s1, const float* s2, const float* lrc,
                     int n, float lambda, int16_t* q4, uint8_t* conf, float* p255);

/* --prior / --c2f (core/tools/de_dense.cpp:797-853, 1156-1296, 1889-1966):
 * the prior as a mask on absolute-disparity planes; sub-pixel off. prior is
 * W*H finite floats. c2f: coarse (nullable) gets the filled W/2 x H/2 coarse
 * disparity, *used_mask 1 when the masked sweep ran (0: holes, full sweep). */
int orc_dense_run_prior(const uint8_t* L, const uint8_t* R, int W, int H, const OrcCfg* cfg,
                        const float* prior, int band, float* disp, float* margin,
                        float* cand6, uint8_t* tie_mask);
void orc_prior_upsample(const float* half, int W, int H, float* full);
int orc_dense_run_c2f(const uint8_t* L, const uint8_t* R, int W, int H, const OrcCfg* cfg,
                      int band, float* disp, float* margin, float* cand6, uint8_t* tie_mask,
                      float* coarse, int* used_mask);

#ifdef __cplusplus
}
#endif
#endif
