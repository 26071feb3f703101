/* C ABI of the MI355X (gfx950) DoubleEye stereo-matching hot path.
 *
 * Plain C: POD structs, pointers and sizes, int status returns (0 = ok), no
 * exceptions across the boundary, caller-owned buffers. Functions named *_dev
 * take DEVICE pointers and an optional hipStream_t (as void*, NULL = the
 * context's stream) and return without synchronising; the others take HOST
 * pointers and return when the results are in host memory.
 *
 * Every entry point cites the reference interface it stands in for (paths
 * relative to the reference repository root). Library: libdoubleeye_hip.so.
 */
#ifndef DOUBLEEYE_HIP_H
#define DOUBLEEYE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DE_OK 0
#define DE_ERR_ARG -1      /* invalid argument (sizes, null pointers, ranges) */
#define DE_ERR_UNSUPPORTED -2
#define DE_ERR_HIP -3      /* HIP runtime error; see de_last_error() */
#define DE_ERR_NOMEM -4
#define DE_ERR_BUSY -5     /* every buffer slot of a serving stream is in use */

/* Dense configuration: the default-path fields of doubleeye::Cfg
 * (core/include/doubleeye/dense_solve.hpp:19-97) plus the census window.
 * census_hw/hh = 3/3 is the reference's 7x7 (48 bits); 4/3 is the 9x7 (62-bit)
 * extension. Disparities searched are d = dmin + k, k in [0, dmax - dmin]. */
typedef struct de_cfg {
  int dmin, dmax, iters;
  int census_hw, census_hh;
  float sigma_s, sigma_r;
  float ad;
  int ad_trunc;
  float lambda, gamma, damping, min_margin, min_ratio;
  int subpixel, fit_eq, lrc;
} de_cfg;

typedef struct de_ctx de_ctx;

/* Reference defaults (dense_solve.hpp:34-97): dmin 1, dmax 60, iters 0,
 * sigma_s 8, sigma_r 0.2, ad 0.15, ad_trunc 10, lambda = gamma = -0.1,
 * damping 0.4, subpixel on, equiangular fit. */
void de_default_cfg(de_cfg* cfg);

/* Context: device, stream, resident scratch (grown on demand, reused). One
 * context may be used from several threads; calls are serialised internally. */
int de_ctx_create(int device, de_ctx** out);
void de_ctx_destroy(de_ctx* ctx);
int de_ctx_sync(de_ctx* ctx);
/* The context's stream (hipStream_t) as void*. */
void* de_ctx_stream(de_ctx* ctx);
/* Per-stage timing with hipEvents recorded between the stage launches on the
 * launching stream (prep, row, column, solve). Enabling resets the sums;
 * de_ctx_stage_ms returns the summed milliseconds per stage and the number of
 * launch sequences measured. No counterpart in the reference, which times its
 * kernels with cudaEvents inside the tool (core/tools/de_dense_cuda.cu:552-565). */
int de_ctx_set_profiling(de_ctx* ctx, int on);
int de_ctx_stage_ms(de_ctx* ctx, double* ms4, int* runs);
/* Launch variants of the two scan kernels (row 0, 1: register-resident pass
 * with and without LDS forward slots; 2, 3: streaming pass with 8 / 16 blocks
 * of backward prefetch. col 0..2: 8/16/32 rows in flight). -1 restores the
 * tuned/default choice. Results never depend on the variant. No counterpart in the
 * reference; it is the knob the autotuner turns. */
int de_ctx_set_variants(de_ctx* ctx, int row_variant, int col_variant);
/* Message of the last failure on this thread ("" if none). */
const char* de_last_error(void);

/* Device memory through the library's own HIP runtime (callers that have no
 * HIP of their own, e.g. ctypes users, the bench and the tests). Copies are on
 * the context's stream and synchronous. */
int de_alloc(de_ctx* ctx, size_t bytes, void** d_ptr);
int de_free(de_ctx* ctx, void* d_ptr);
int de_upload(de_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int de_download(de_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
/* Plain device copy / byte fill on the context's stream (asynchronous). */
int de_copy_dev(de_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
int de_fill_dev(de_ctx* ctx, void* d_dst, int byte, size_t bytes);

/* Census transform of one image, bit b in dy-major/dx-minor order with the
 * centre skipped, set iff neighbour < centre, 0 where the window is cut.
 * Replaces census_transform (core/include/doubleeye/preproc.hpp:92,
 * core/src/preproc.cpp:110-138) and census_image (core/tools/de_dense.cpp:186).
 * hw/hh in {1/1, 2/2, 3/3, 4/3}. */
int de_census(de_ctx* ctx, const uint8_t* img, int W, int H, int hw, int hh,
              uint64_t* out);
int de_census_dev(de_ctx* ctx, const uint8_t* d_img, int W, int H, int hw, int hh,
                  uint64_t* d_out, void* stream);

/* The dense pipeline of run_dense (core/tools/de_dense.cpp:738-1608) on its
 * default path (recursive filter, blockwise top-2, K = 2 row solve), equal to
 * `de_dense --threads 1`. Outputs: disp W*H float (NaN = no answer),
 * margin W*H float, cand6 (nullable) 6 planes s1, s2, d1, d2, lrc, lrc2 as
 * written by --out-conf (-1e30 = missing candidate). Rows up to 4096 pixels
 * (the row solver holds a row in LDS); wider images return DE_ERR_ARG. */
int de_dense_run(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W, int H,
                 const de_cfg* cfg, float* disp, float* margin, float* cand6);

/* Pinned host memory for the host-pointer entry points. The reference's
 * run_dense takes ordinary heap buffers (core/tools/de_dense.cpp:738-741:
 * Image8 pair, std::vector output planes, loaded at :1751); every host-pointer
 * entry point accepts those, staging them through the context's own pinned
 * buffers. Buffers from
 * de_host_alloc are read and written by DMA directly, with no host copy.
 * de_host_free takes only what de_host_alloc returned. */
int de_host_alloc(size_t bytes, void** out);
int de_host_free(void* p);
int de_dense_run_dev(de_ctx* ctx, const uint8_t* d_L, const uint8_t* d_R, int W,
                     int H, const de_cfg* cfg, float* d_disp, float* d_margin,
                     float* d_cand6, void* stream);

/* de_dense_run_dev without any device allocation. No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The frame's scratch,
 * tables and launch descriptor live in one static device array of 256 MiB per
 * device (held from library load to process end; its address is looked up
 * when a context is created on the device), filled by kernels from their
 * arguments: a call issues kernel launches only. Frames whose scratch exceeds
 * de_dense_noalloc_limit() bytes (848x480 at 128 disparities fits) are refused
 * with DE_ERR_ARG. Every call on a device shares the array: calls that overlap
 * on one device (other host threads, unordered streams) race. Results equal
 * de_dense_run_dev's. */
int de_dense_run_dev_noalloc(de_ctx* ctx, const uint8_t* d_L, const uint8_t* d_R, int W, int H,
                             const de_cfg* cfg, float* d_disp, float* d_margin, float* d_cand6,
                             void* stream);
size_t de_dense_noalloc_limit(void);

/* The FP32 full-volume path (`de_dense --no-blockwise/--dump-vol`): the float
 * score and recursive filter (core/tools/de_dense.cpp:379-396, 510-527,
 * 1305-1318), volume [y][x][k] of W*H*(dmax-dmin+1) floats, -1e30 outside each
 * pixel's valid window, exactly the --dump-vol file (de_dense.cpp:1433-1436).
 * Host pointers. */
int de_dense_volume_f32(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W, int H,
                        const de_cfg* cfg, float* vol);

/* `de_dense --topk 0`: the FP32 volume and the full-volume row solver
 * solve_row (de_dense.cpp:566-712). disp (NaN = no answer) and margin, W*H. */
int de_dense_run_full(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W, int H,
                      const de_cfg* cfg, float* disp, float* margin);
int de_dense_run_full_dev(de_ctx* ctx, const uint8_t* d_L, const uint8_t* d_R, int W, int H,
                          const de_cfg* cfg, float* d_disp, float* d_margin, void* stream);

/* ---- Batches: many pairs of mixed size and disparity range, one launch per
 * stage (plus one column-pass launch per disparity-group count). No
 * counterpart in the reference, which runs one fixed-geometry pair at a time
 * (core/tools/de_dense_cuda.cu:670-770). */

/* One pair of a batch. Device pointers; conf and q4 (nullable) are the DEDD
 * output planes of core/tools/de_dense_cuda.cu:731-745 (u8 confidence, Q4
 * disparity with -32768 = no answer). */
typedef struct de_pair {
  const uint8_t* L;
  const uint8_t* R;
  int W, H;
  float* disp;
  float* margin;
  float* cand6;
  uint8_t* conf;
  int16_t* q4;
} de_pair;

/* Run n pairs with per-pair configurations (D, dmin, iters and the gates may
 * differ; the census window, sigma_s, sigma_r, ad and ad_trunc must be equal).
 * Results are identical to n separate de_dense_run_dev calls. */
int de_dense_batch_dev(de_ctx* ctx, const de_pair* pairs, const de_cfg* cfgs, int n,
                       void* stream);

/* A captured batch: its own resident scratch, descriptors uploaded once and
 * the stage launches recorded in hipGraphs. Launching replays them on
 * `stream`; the caller refreshes the pairs' input buffers before each launch
 * (e.g. hipMemcpyAsync on the same stream). Plans of one context may be in
 * flight concurrently on different streams; one plan runs one launch at a
 * time (launches queued on one stream follow each other).
 * With DE_PLAN_PARTS = k (2..8) in the environment at creation, the batch is
 * split into k parts of equal px*disp work: part i's prep and row pass run on
 * one internal stream in part order, its column pass and solver on a second
 * one after its own row pass, so one part's column pass overlaps the next
 * part's row pass. The default is one graph of every stage. The results are
 * the same bit for bit either way.
 * (No reference counterpart: the reference tool runs one pair at a time,
 * core/tools/de_dense_cuda.cu:670-770.) */
typedef struct de_plan de_plan;
int de_plan_create(de_ctx* ctx, const de_pair* pairs, const de_cfg* cfgs, int n,
                   de_plan** out);
int de_plan_launch(de_plan* plan, void* stream);
void de_plan_destroy(de_plan* plan);
/* Number of parts of a plan (1: one graph of every stage). */
int de_plan_parts(const de_plan* plan);

/* Autotuner: times every prefetch-depth variant of the row and column passes
 * on a synthetic pair of this shape (reps launches each), remembers the
 * fastest for the shape bucket (W, H, disparity-group count) and returns it.
 * Later launches of that bucket (plans included) use it. */
int de_autotune(de_ctx* ctx, int W, int H, const de_cfg* cfg, int reps, int* row_variant,
                int* col_variant);

/* Intermediate for parity tests: the aggregated int16 Q14 volume for
 * disparities [k0, k1), layout [k][y][x] (the plane rf_filter_i16 leaves,
 * core/tools/de_dense.cpp:485-508). Host pointers. */
int de_dense_volume(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W, int H,
                    const de_cfg* cfg, int k0, int k1, int16_t* out);

/* Intermediate for parity tests: per-pixel candidate records after the fused
 * vertical filter + top-2 (4 x uint32: key0, key1, neighbour pair, 0), host
 * pointer W*H*4 words. key = ((v + 16385) << 8) | (255 - k), 0 = none. */
int de_dense_candidates(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W,
                        int H, const de_cfg* cfg, uint32_t* out);

/* ==== Sparse keypoint MASDA (core/src/match.cpp) ==========================
 * Left keypoints are measurements, right keypoints objects; candidate edges
 * join a left keypoint to right keypoints in its y band and disparity window,
 * scored by Census agreement and the y residual (plus optional soft priors).
 * Results equal the reference's whenever no exact tie decides an order; at
 * exact ties every order is broken by position (right keypoints by (y, index),
 * kept candidates by (score, band position), greedy completion by (belief,
 * score, edge index)), where the reference follows libstdc++'s unstable sorts. */

/* doubleeye::Keypoint (core/include/doubleeye/preproc.hpp:125-130) */
typedef struct de_keypoint {
  float x, y, response, local_std;
} de_keypoint;

/* doubleeye::Match (core/include/doubleeye/match.hpp:46-65) */
typedef struct de_match {
  int left, right;
  float score, belief, disparity, dy, margin;
} de_match;

/* doubleeye::MatchConfig (match.hpp:136-176); its prior and smooth pointers
 * are the separate arguments below. */
typedef struct de_match_cfg {
  float max_dy, min_disparity, max_disparity;
  int max_candidates;
  float sigma_y, w_y, w_prior, w_smooth;
  int iterations;
  float damping, lambda, gamma, converge_eps;
} de_match_cfg;

/* doubleeye::DisparityPrior (match.hpp:81-92): cols*rows cells of `cell` px. */
typedef struct de_disparity_prior {
  int cell, cols, rows;
  const float* disparity;
  const float* sigma;
  const uint8_t* valid;
} de_disparity_prior;

/* doubleeye::SmoothPrior (match.hpp:110-114): per left keypoint. */
typedef struct de_smooth_prior {
  int n;
  const float* disparity;
  const float* sigma;
  const uint8_t* valid;
} de_smooth_prior;

/* doubleeye::MatchStats (match.hpp:178-186) */
typedef struct de_match_stats {
  int left, right, candidates, iterations_run, converged;
  float final_max_delta;
  int oscillating;
} de_match_stats;

/* Right keypoints per frame (the y sort runs in one workgroup's LDS). */
#define DE_SPARSE_MAX_RIGHT 8192

/* MatchConfig's defaults (match.hpp:136-176). */
void de_match_default_cfg(de_match_cfg* cfg);

/* match_masda (match.hpp:221-226, match.cpp:334-417). Host pointers; prior and
 * smooth are nullable. `out` holds min(n_left, n_right) matches; *n_out gets
 * the count; stats is nullable. */
int de_match_masda(de_ctx* ctx, const de_keypoint* left, const uint64_t* desc_left,
                   int n_left, const de_keypoint* right, const uint64_t* desc_right,
                   int n_right, const de_match_cfg* cfg, const de_disparity_prior* prior,
                   const de_smooth_prior* smooth, de_match* out, int* n_out,
                   de_match_stats* stats);

/* match_mutual_nn (match.hpp:228-235, match.cpp:419-476): mutual nearest
 * neighbour with a ratio test over the same candidate graph. */
int de_match_mutual_nn(de_ctx* ctx, const de_keypoint* left, const uint64_t* desc_left,
                       int n_left, const de_keypoint* right, const uint64_t* desc_right,
                       int n_right, const de_match_cfg* cfg,
                       const de_disparity_prior* prior, const de_smooth_prior* smooth,
                       float ratio, de_match* out, int* n_out, de_match_stats* stats);

/* The candidate graph build_candidates (match.cpp:46-102) produces, in the
 * solver's edge order: edge e joins ei[e] to ej[e] with score es[e]. Arrays
 * hold n_left * max(1, max_candidates) entries. For parity tests. */
int de_match_candidates(de_ctx* ctx, const de_keypoint* left, const uint64_t* desc_left,
                        int n_left, const de_keypoint* right, const uint64_t* desc_right,
                        int n_right, const de_match_cfg* cfg,
                        const de_disparity_prior* prior, const de_smooth_prior* smooth,
                        int* ei, int* ej, float* es, int* n_edges);

/* One sparse frame of a device batch. Device pointers; out holds
 * min(n_left, n_right) matches; n_out and stats are single device records. */
typedef struct de_sparse_frame {
  const de_keypoint* left;
  const uint64_t* desc_left;
  int n_left;
  const de_keypoint* right;
  const uint64_t* desc_right;
  int n_right;
  de_match* out;
  int* n_out;
  de_match_stats* stats;
} de_sparse_frame;

/* match_masda over n frames with per-frame configurations, three launches in
 * all (y sort, candidates, solve: one workgroup per frame). No priors. Results
 * equal n separate de_match_masda calls. No counterpart in the reference. */
int de_match_masda_batch_dev(de_ctx* ctx, const de_sparse_frame* frames,
                             const de_match_cfg* cfgs, int n, void* stream);

/* ==== The prior as a mask (--prior, --c2f) =================================
 * core/tools/de_dense.cpp:797-853, 1156-1296, 1771-1966. Plane d is evaluated
 * on the rectangle bounding the pixels whose band |d - round(prior)| <= band
 * admits it (padded by 8, clipped to the score region); the recursive filter
 * runs on that rectangle only, from its own edge values, and a score becomes a
 * candidate only inside its row's interval. Sub-pixel is off (the reference
 * turns it off under a mask). Outputs as de_dense_run. */

/* --prior: `prior` holds W*H values, or W/2 * H/2 (upsampled nearest, the
 * disparity doubled, as the reference does for a half-resolution prior); every
 * value finite (DE_ERR_ARG otherwise). band: the reference's --band (2). */
int de_dense_run_prior(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W, int H,
                       const de_cfg* cfg, const float* prior, size_t n_prior, int band,
                       float* disp, float* margin, float* cand6);

/* --c2f: the half-resolution pass (dmin/2 .. (dmax+1)/2, one iteration,
 * min_margin 0), hole fill, upsample, then the masked sweep; the full sweep
 * when holes remain. coarse (nullable) gets the filled W/2 x H/2 coarse map,
 * *used_mask (nullable) 1 when the masked sweep ran. */
int de_dense_run_c2f(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W, int H,
                     const de_cfg* cfg, int band, float* disp, float* margin, float* cand6,
                     float* coarse, int* used_mask);

/* The same (core/tools/de_dense.cpp:1771-1966) on device pointers, queued on
 * `stream` (null: the context's). The coarse level's hole count decides
 * between the masked and the full sweep, so the call waits for the stream
 * once, after the coarse level; it returns with the rest queued. */
int de_dense_run_c2f_dev(de_ctx* ctx, const uint8_t* d_L, const uint8_t* d_R, int W, int H,
                         const de_cfg* cfg, int band, float* d_disp, float* d_margin,
                         float* d_cand6, int* used_mask, void* stream);

/* ==== Keypoint detector and the detect -> match pipeline ===================
 * detect_keypoints_fast (core/src/preproc.cpp:436-545): FAST-9 candidates,
 * Shi-Tomasi response at candidates, greedy suppression within nms_radius in
 * descending response order, texture floor, sub-pixel parabola, the top
 * per_cell per cell x cell block; keypoints in row-major cell order, each
 * cell's in descending response. Equal responses are ordered by scan position
 * (y, then x), where the reference's std::sort leaves their order open. */

/* doubleeye::DetectorConfig (core/include/doubleeye/preproc.hpp:132-145);
 * subpixel is 0 or 1. */
typedef struct de_detector_cfg {
  int grad_window, cell, per_cell;
  float min_response, min_local_std;
  int border, subpixel, fast_threshold, nms_radius;
} de_detector_cfg;

/* doubleeye::DetectProfile (preproc.hpp:191-198): stage times from GPU events. */
typedef struct de_detect_profile {
  double fast_ms, nms_ms, refine_ms;
  int candidates, suppressed;
} de_detect_profile;

/* One image of a device detector batch. Device pointers; cap >=
 * de_detect_max_keypoints(W, H, cfg); n_out is a single device int. */
typedef struct de_detect_frame {
  const uint8_t* img;
  int W, H;
  de_keypoint* out;
  int cap;
  int* n_out;
} de_detect_frame;

/* DetectorConfig's defaults. */
void de_detector_default_cfg(de_detector_cfg* cfg);

/* Upper bound on the keypoints of one image: min(cells * per_cell, pixels
 * inside the border). */
size_t de_detect_max_keypoints(int W, int H, const de_detector_cfg* cfg);

/* detect_keypoints_fast (preproc.hpp:202-204). Host pointers; fails with
 * DE_ERR_ARG when more than `cap` keypoints result (*n_out then holds the
 * count). prof is nullable; its stage times come from GPU events.
 * nms_radius is limited to 255. */
int de_detect_keypoints_fast(de_ctx* ctx, const uint8_t* img, int W, int H,
                             const de_detector_cfg* cfg, de_keypoint* out, int cap, int* n_out,
                             de_detect_profile* prof);

/* The detector over n images with per-image configurations, in one chain of
 * launches on `stream` without host synchronisation. Results equal n separate
 * de_detect_keypoints_fast calls. No counterpart in the reference. */
int de_detect_keypoints_fast_batch_dev(de_ctx* ctx, const de_detect_frame* frames,
                                       const de_detector_cfg* cfgs, int n, void* stream);

/* is_fast_corner (preproc.hpp:179) and shi_tomasi_at (preproc.hpp:183) at
 * every pixel: corner[y*W+x] in {0, 1}, response[y*W+x]. For parity tests. */
int de_fast_score_map(de_ctx* ctx, const uint8_t* img, int W, int H, int threshold,
                      int grad_window, uint8_t* corner, float* response);

/* census_at (preproc.hpp:96-97) at (int(x + 0.5f), int(y + 0.5f)) of each
 * keypoint, as core/tools/de_pipe.cpp:168-172 describes them. */
int de_census_at_keypoints(de_ctx* ctx, const uint8_t* img, int W, int H, int hw, int hh,
                           const de_keypoint* kp, int n, uint64_t* desc);

/* refine_disparity (match.hpp:293-296, match.cpp:768-810): in place on the
 * host matches; half is the SAD half-window (the reference's default is 3). */
int de_refine_disparity(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W, int H,
                        const de_keypoint* kl, int n_left, const de_keypoint* kr, int n_right,
                        de_match* matches, int n, int half);

/* One stereo pair through core/tools/de_pipe.cpp:167-199: detect on both images
 * (det_right nullable = det_left), census_at at the keypoints, match_masda,
 * drop matches whose margin is below min_margin (when > 0), refine_disparity
 * (half 3) when subpixel. Outputs the keypoints, the matches and, when
 * `records` is non-null, the DEMR records (x_left, y_left, disparity, margin)
 * as 4 floats per match. Capacities: kl_cap / kr_cap from
 * de_detect_max_keypoints, out_cap >= min(n_left, n_right). */
int de_sparse_pipeline(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W, int H,
                       const de_detector_cfg* det_left, const de_detector_cfg* det_right,
                       int census_hw, int census_hh, const de_match_cfg* mcfg, float min_margin,
                       int subpixel, de_keypoint* kl, int kl_cap, int* n_kl, de_keypoint* kr,
                       int kr_cap, int* n_kr, de_match* out, float* records, int out_cap,
                       int* n_out, de_match_stats* stats);

/* ==== Keypoint-graph belief propagation ====================================
 * No counterpart in the reference (its sparse matcher is bipartite). For
 * keypoints of the left image: labels are the n_cand disparities along the
 * same row with the smallest census Hamming distance (ties by smaller d), unary
 * cost w_data * Hamming; the graph joins each keypoint to its k_nn nearest
 * neighbours (symmetric union); min-sum loopy BP with the smoothness cost
 * lambda * min(|d_i - d_j|, trunc) runs `iters` synchronous updates from zero
 * messages; the output is the best and second label per keypoint. All float
 * arithmetic follows one fixed order, so results are reproducible bit for bit. */
typedef struct de_knn_cfg {
  int census_hw, census_hh;  /* 4, 3: the 9x7 window */
  int dmax;                  /* disparities 0..dmax, dmax <= 255 */
  int n_cand;                /* labels per keypoint, 1..8 */
  int k_nn;                  /* neighbours per keypoint, 0..8 */
  int iters;
  float w_data, lambda;
  int trunc;                 /* px */
} de_knn_cfg;

typedef struct de_knn_out {
  int d1, d2;                /* best and second disparity, -1 when absent */
  float c1, c2;              /* their belief costs, +inf when absent */
} de_knn_out;

void de_knn_default_cfg(de_knn_cfg* cfg);

/* The whole workload on one frame. Host pointers; kp is n (x, y) int pairs,
 * each inside the census border. */
int de_knn_bp(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W, int H,
              const int32_t* kp, int n, const de_knn_cfg* cfg, de_knn_out* out);

/* Intermediates for parity tests (host pointers): labels [n][8] (d = -1 and
 * cost 0 past ncand[i]), and the graph as CSR (off [n+1], adj ascending). */
int de_knn_candidates(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W, int H,
                      const int32_t* kp, int n, const de_knn_cfg* cfg, int32_t* cand_d,
                      float* cand_c, int32_t* ncand);
int de_knn_graph(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W, int H,
                 const int32_t* kp, int n, const de_knn_cfg* cfg, int32_t* off,
                 int32_t* adj, int adj_cap, int* n_edges);

/* BP alone on a given symmetric graph (ascending CSR lists, at most 64
 * neighbours and 8 labels per node). beliefs is nullable, [n][8]. */
int de_knn_bp_graph(de_ctx* ctx, int n, const int32_t* cand_d, const float* cand_c,
                    const int32_t* ncand, const int32_t* off, const int32_t* adj,
                    const de_knn_cfg* cfg, de_knn_out* out, float* beliefs);

/* One frame of a device batch (device pointers). */
typedef struct de_knn_frame {
  const uint8_t* L;
  const uint8_t* R;
  int W, H;
  const int32_t* kp;
  int n;
  de_knn_out* out;
} de_knn_frame;

/* The workload over n frames (equal iters), one launch per stage. Keypoints
 * outside the census border get no labels (d1 = -1). */
int de_knn_bp_batch_dev(de_ctx* ctx, const de_knn_frame* frames, const de_knn_cfg* cfgs,
                        int n, void* stream);

/* ==== Edge-list MASDA in float64: the NumPy call shapes =====================
 * article/masda_stereo.py and article/dense_masda.py. Messages and beliefs in
 * float64 with numpy's rounding; the greedy decision takes edges in (belief
 * desc, edge index asc) order, where numpy's unstable argsort may order exact
 * ties differently. Host pointers; index arrays are int64 like numpy's. */

/* dense_problem (dense_masda.py:39-79): every interior pixel against every
 * candidate d in [dmin, dmax] on its row; edges d-major, then y, then x; left
 * index y*W + x, right index y*W + x - d, score (24 - Hamming) / 24 (float32).
 * With ei, ej and se all NULL only *n_edges is computed. The node count is W*H. */
int de_dense_problem(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W, int H,
                     int dmin, int dmax, int census_hw, int census_hh, int64_t* ei,
                     int64_t* ej, float* se, int64_t capacity, int64_t* n_edges);

/* masda_sparse (masda_stereo.py:453-492). assign[i] = j or -1 for the m left
 * nodes; hist gets the largest message change of each iteration run
 * (*n_iter of them, at most iters). The _f32 form takes float32 scores (as
 * dense_problem returns them); they enter the float64 arithmetic exactly. */
int de_masda_edges(de_ctx* ctx, const int64_t* ei, const int64_t* ej, const double* se,
                   int64_t n_edges, int64_t m, int64_t n, double lam, double gam, int iters,
                   double damping, double eps, int64_t* assign, double* hist, int* n_iter);
int de_masda_edges_f32(de_ctx* ctx, const int64_t* ei, const int64_t* ej, const float* se,
                       int64_t n_edges, int64_t m, int64_t n, double lam, double gam,
                       int iters, double damping, double eps, int64_t* assign, double* hist,
                       int* n_iter);

/* masda (masda_stereo.py:329-389): a dense m x n float32 score matrix, -inf
 * where there is no candidate; same solver on its row-major edges. */
int de_masda_matrix(de_ctx* ctx, const float* S, int64_t m, int64_t n, double lam,
                    double gam, int iters, double damping, double eps, int64_t* assign,
                    double* hist, int* n_iter);

/* decide (masda_stereo.py:392-414) on an edge list: greedy in belief order,
 * an edge eligible when its score beats lam and its belief is finite. */
int de_masda_decide(de_ctx* ctx, const int64_t* ei, const int64_t* ej, const double* se,
                    const double* belief, int64_t n_edges, int64_t m, int64_t n, double lam,
                    int64_t* assign);

/* _seg_max_excluding (masda_stereo.py:417-442): out[e] = max over the other
 * elements of segment idx[e] (-inf when it is alone). */
int de_seg_max_excluding(de_ctx* ctx, const double* vals, const int64_t* idx, int64_t n_vals,
                         int64_t n_segments, double* out);

/* census(img, pts, half_w, half_h) (masda_stereo.py:264-283): out[i] = the
 * census descriptor at (xy[2i], xy[2i+1]) (integer pixel positions; a negative
 * index counts from the end, as numpy indexing does). Refuses a point whose
 * window reaches outside [-W, W) x [-H, H) (numpy raises there) and windows
 * of more than 64 bits. */
int de_census_points(de_ctx* ctx, const uint8_t* img, int W, int H, const int64_t* xy,
                     int64_t n_points, int half_w, int half_h, uint64_t* out);

/* hamming(a, b) (masda_stereo.py:289-294): out[i] = popcount(a[i] ^ b[i]). */
int de_hamming(de_ctx* ctx, const uint64_t* a, const uint64_t* b, int64_t n, int32_t* out);

/* build_problem (masda_stereo.py:305-326): the m x n score matrix S, row-major,
 * S[i][j] = (24 - Hamming(dl[i], dr[j])) / 24 - w_y (dy / sigma_y)^2 where
 * |dy| <= max_dy and dmin <= d <= dmax (dy, d: left minus right y, x), -inf
 * elsewhere. pl, pr: (x, y) pairs, float32 (fp64 = 0) or float64 (fp64 = 1);
 * the arithmetic follows numpy's types for each (the parameters take the
 * points' type). */
int de_build_problem(de_ctx* ctx, const void* pl, const uint64_t* dl, int64_t m, const void* pr,
                     const uint64_t* dr, int64_t n, int fp64, double max_dy, double dmin,
                     double dmax, double sigma_y, double w_y, float* S);

/* dense_problem + masda_sparse + the disparity map of dense_masda.run
 * (dense_masda.py:82-110, without the optional margin gate), all on the
 * device: disp[y*W + x] = x - (j mod W) for matched pixels, NaN elsewhere. */
int de_dense_masda(de_ctx* ctx, const uint8_t* L, const uint8_t* R, int W, int H, int dmin,
                   int dmax, int census_hw, int census_hh, double lam, double gam, int iters,
                   double damping, double eps, float* disp, double* hist, int* n_iter);

/* ==== DEIR -> DEDD stream front end ========================================
 * The packet grammar and pairing of `de_dense_cuda --stream`
 * (core/tools/de_dense_cuda.cu:44-59, 670-770):
 *   in:  "DEIR" | w u16 | h u16 | stream u8 | flags u8 | frame u32 | w*h Y8
 *   out: "DEDD" | w u16 | h u16 | frame u32 | flags u32 | left Y8 |
 *        w*h int16 Q4 disparity (-32768 = none) | w*h u8 confidence if flags & 1
 * Images wait by frame number (8 at most, the oldest dropped beyond that); a
 * second image of the same stream replaces the first; stream 1 is the left
 * camera. A bad magic or geometry returns DE_ERR_ARG (the reference exits 1). */
typedef struct de_stream de_stream;
/* Output bytes of one DEDD packet. */
size_t de_stream_packet_bytes(int W, int H, int want_conf);
int de_stream_create(de_ctx* ctx, int W, int H, const de_cfg* cfg, int want_conf,
                     de_stream** out);
void de_stream_destroy(de_stream* s);
/* Feed any number of input bytes; the DEDD packets of every pair they
 * complete are written to out (capacity out_cap; de_stream_packet_bytes per
 * pair completed), *out_len bytes. Up to three pairs run at once inside one
 * call (own streams and plans); packets keep pairing order. Packets past
 * out_cap are dropped and the call returns DE_ERR_ARG. */
int de_stream_feed(de_stream* s, const uint8_t* data, size_t n, uint8_t* out, size_t out_cap,
                   size_t* out_len);
int de_stream_stats(de_stream* s, long* pairs, long* dropped);
/* The whole loop over file descriptors until end of input. Pairs stay in
 * flight while more input is already readable; packets are written before
 * the loop blocks on a read, and at the end. */
int de_stream_run_fd(de_ctx* ctx, int in_fd, int out_fd, int W, int H, const de_cfg* cfg,
                     int want_conf, long* pairs, long* dropped);

/* ==== Serving loop =========================================================
 * Many camera streams of mixed geometry on one GPU, with continuous batching.
 * The reference serves one fixed-geometry stream with one frame in flight
 * (`de_dense_cuda --stream`, core/tools/de_dense_cuda.cu:670-770, driven by
 * desktop/de_live_ros2.py). A stream is opened against the server's device
 * memory budget (its scratch, sized from W x H x D or W x H x keypoints, plus
 * `slots` input/output buffer sets, all resident for the stream's life).
 * Submitting copies a frame into a free slot; a scheduler thread batches the
 * oldest waiting frame of every stream whose previous frame has finished and
 * launches the batch (copies included) on an idle HIP stream, so batches
 * overlap; results wait in pinned host memory until released. A frame's
 * result equals a plain call on the same inputs. Dense streams of one server
 * share the census window, sigma_s, sigma_r, ad and ad_trunc; while a server
 * runs, dense calls on its context must use the same values. */
typedef struct de_server de_server;

typedef struct de_server_cfg {
  size_t mem_budget;   /* device bytes all open streams may hold together */
  int hip_streams;     /* batches in flight at once, 1..8 */
  int slots;           /* buffer sets per camera stream, 1..16 */
  int max_batch;       /* frames per batch, 0 = no limit */
} de_server_cfg;

#define DE_STREAM_DENSE 0
#define DE_STREAM_KNN 1

typedef struct de_server_result {
  int stream_id, slot, kind, status;
  uint64_t tag;
  double t_submit, t_start, t_done;   /* seconds, monotonic clock */
  int batch_size;                     /* frames launched together */
  int W, H, n;                        /* geometry; keypoints (knn) */
  /* host results, valid until de_server_release(stream_id, slot):
   * dense: disp W*H (NaN = none), q4 W*H (-32768 = none), conf W*H;
   * knn: knn[n] */
  const float* disp;
  const int16_t* q4;
  const uint8_t* conf;
  const de_knn_out* knn;
} de_server_result;

/* 8 GiB, 4 HIP streams, 4 slots, no batch limit. */
void de_server_default_cfg(de_server_cfg* cfg);
int de_server_create(de_ctx* ctx, const de_server_cfg* cfg, de_server** out);
/* Waits for the batches in flight; waiting frames are dropped. */
void de_server_destroy(de_server* server);
/* DE_ERR_NOMEM when the stream does not fit the remaining budget. */
int de_server_open_dense(de_server* server, int W, int H, const de_cfg* cfg, int* stream_id);
int de_server_open_knn(de_server* server, int W, int H, int max_kp, const de_knn_cfg* cfg,
                       int* stream_id);
/* Waits for the stream's frames in flight, drops its waiting frames and
 * results, returns its memory to the budget. */
int de_server_close(de_server* server, int stream_id);
int de_server_memory(de_server* server, size_t* used, size_t* budget);
/* Host pointers, copied before return. kp/n_kp: knn streams only.
 * DE_ERR_BUSY when every slot of the stream holds a frame or a result. */
int de_server_submit(de_server* server, int stream_id, const uint8_t* L, const uint8_t* R,
                     const int32_t* kp, int n_kp, uint64_t tag, int* slot);
/* Up to max finished frames, oldest first; waits up to timeout_ms (0: no
 * wait, < 0: until one is ready). */
int de_server_poll(de_server* server, de_server_result* results, int max, int timeout_ms,
                   int* n);
int de_server_release(de_server* server, int stream_id, int slot);

#ifdef __cplusplus
}
#endif

#endif /* DOUBLEEYE_HIP_H */
