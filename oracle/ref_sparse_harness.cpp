// C entry points over the reference's own sparse matcher, compiled from
// /root/reference/core/src/match.cpp (+ preproc.cpp) by oracle/Makefile.ref into
// oracle/_ref/libref_sparse.so. TEST INFRASTRUCTURE ONLY: it pins the oracle
// restatement in oracle/sparse_ref.cpp and writes the golden fixtures.
//
// The reference API used here is doubleeye::match_masda / match_mutual_nn /
// match_brute_force (core/include/doubleeye/match.hpp:221-261).
#include "doubleeye/match.hpp"

#include <cstring>
#include <vector>

#include "sparse_ref.h"

namespace {

using namespace doubleeye;

std::vector<Keypoint> kps(const OrcKeypoint* p, int n) {
  std::vector<Keypoint> v(size_t(n > 0 ? n : 0));
  for (int k = 0; k < n; ++k) {
    v[size_t(k)].x = p[k].x;
    v[size_t(k)].y = p[k].y;
    v[size_t(k)].response = p[k].response;
    v[size_t(k)].local_std = p[k].local_std;
  }
  return v;
}

std::vector<uint64_t> descs(const uint64_t* p, int n) {
  return std::vector<uint64_t>(p, p + (n > 0 ? n : 0));
}

struct Cfg {
  MatchConfig c;
  DisparityPrior prior;
  SmoothPrior smooth;
};

void to_cfg(const OrcMatchCfg* o, const OrcPrior* pr, const OrcSmooth* sm, Cfg* out) {
  MatchConfig& c = out->c;
  c.max_dy = o->max_dy;
  c.min_disparity = o->min_disparity;
  c.max_disparity = o->max_disparity;
  c.max_candidates = o->max_candidates;
  c.sigma_y = o->sigma_y;
  c.w_y = o->w_y;
  c.w_prior = o->w_prior;
  c.w_smooth = o->w_smooth;
  c.iterations = o->iterations;
  c.damping = o->damping;
  c.lambda = o->lambda;
  c.gamma = o->gamma;
  c.converge_eps = o->converge_eps;
  c.prior = nullptr;
  c.smooth = nullptr;
  if (pr) {
    const size_t n = size_t(pr->cols) * size_t(pr->rows);
    out->prior.cell = pr->cell;
    out->prior.cols = pr->cols;
    out->prior.rows = pr->rows;
    out->prior.disparity.assign(pr->disparity, pr->disparity + n);
    out->prior.sigma.assign(pr->sigma, pr->sigma + n);
    out->prior.valid.assign(pr->valid, pr->valid + n);
    c.prior = &out->prior;
  }
  if (sm) {
    out->smooth.disparity.assign(sm->disparity, sm->disparity + sm->n);
    out->smooth.sigma.assign(sm->sigma, sm->sigma + sm->n);
    out->smooth.valid.assign(sm->valid, sm->valid + sm->n);
    c.smooth = &out->smooth;
  }
}

void put(const std::vector<Match>& m, OrcMatch* out, int* n_out) {
  for (size_t k = 0; k < m.size(); ++k) {
    out[k].left = m[k].left;
    out[k].right = m[k].right;
    out[k].score = m[k].score;
    out[k].belief = m[k].belief;
    out[k].disparity = m[k].disparity;
    out[k].dy = m[k].dy;
    out[k].margin = m[k].margin;
  }
  *n_out = int(m.size());
}

void put_stats(const MatchStats& s, OrcMatchStats* o) {
  if (!o) return;
  o->left = s.left;
  o->right = s.right;
  o->candidates = s.candidates;
  o->iterations_run = s.iterations_run;
  o->converged = s.converged ? 1 : 0;
  o->final_max_delta = s.final_max_delta;
  o->oscillating = s.oscillating;
}

}  // namespace

extern "C" {

int ref_match_masda(const OrcKeypoint* L, const uint64_t* dl, int nl,
                    const OrcKeypoint* R, const uint64_t* dr, int nr,
                    const OrcMatchCfg* cfg, const OrcPrior* prior,
                    const OrcSmooth* smooth, OrcMatch* out, int* n_out,
                    OrcMatchStats* stats) {
  Cfg c;
  to_cfg(cfg, prior, smooth, &c);
  MatchStats st;
  put(match_masda(kps(L, nl), descs(dl, nl), kps(R, nr), descs(dr, nr), c.c, &st),
      out, n_out);
  put_stats(st, stats);
  return 0;
}

int ref_match_mutual_nn(const OrcKeypoint* L, const uint64_t* dl, int nl,
                        const OrcKeypoint* R, const uint64_t* dr, int nr,
                        const OrcMatchCfg* cfg, const OrcPrior* prior,
                        const OrcSmooth* smooth, float ratio, OrcMatch* out,
                        int* n_out, OrcMatchStats* stats) {
  Cfg c;
  to_cfg(cfg, prior, smooth, &c);
  MatchStats st;
  put(match_mutual_nn(kps(L, nl), descs(dl, nl), kps(R, nr), descs(dr, nr), c.c,
                      ratio, &st),
      out, n_out);
  put_stats(st, stats);
  return 0;
}

int ref_match_brute_force(const OrcKeypoint* L, const uint64_t* dl, int nl,
                          const OrcKeypoint* R, const uint64_t* dr, int nr,
                          const OrcMatchCfg* cfg, OrcMatch* out, int* n_out) {
  Cfg c;
  to_cfg(cfg, nullptr, nullptr, &c);
  put(match_brute_force(kps(L, nl), descs(dl, nl), kps(R, nr), descs(dr, nr), c.c),
      out, n_out);
  return 0;
}

// detect_keypoints_fast (core/include/doubleeye/preproc.hpp:202-204) on a
// W*H Y8 image. Config fields in DetectorConfig order; counts (nullable) gets
// DetectProfile's candidates and suppressed.
int ref_detect_keypoints_fast(const uint8_t* img, int W, int H, const int* icfg,
                              const float* fcfg, OrcKeypoint* out, int cap, int* n_out,
                              int* counts) {
  Image8 im(W, H);
  std::memcpy(im.data.data(), img, size_t(W) * H);
  DetectorConfig c;
  c.grad_window = icfg[0];
  c.cell = icfg[1];
  c.per_cell = icfg[2];
  c.min_response = fcfg[0];
  c.min_local_std = fcfg[1];
  c.border = icfg[3];
  c.subpixel = icfg[4] != 0;
  c.fast_threshold = icfg[5];
  c.nms_radius = icfg[6];
  DetectProfile prof;
  const std::vector<Keypoint> k = detect_keypoints_fast(im, c, &prof);
  if (counts) {
    counts[0] = prof.candidates;
    counts[1] = prof.suppressed;
  }
  if (int(k.size()) > cap) return -1;
  for (size_t i = 0; i < k.size(); ++i) {
    out[i].x = k[i].x;
    out[i].y = k[i].y;
    out[i].response = k[i].response;
    out[i].local_std = k[i].local_std;
  }
  *n_out = int(k.size());
  return 0;
}

int ref_is_fast_corner(const uint8_t* img, int W, int H, int x, int y, int thr) {
  Image8 im(W, H);
  std::memcpy(im.data.data(), img, size_t(W) * H);
  return is_fast_corner(im, x, y, thr) ? 1 : 0;
}

// shi_tomasi_at over every pixel (preproc.hpp:183), and the dense
// shi_tomasi_response (preproc.hpp:151) for comparison.
void ref_shi_tomasi_maps(const uint8_t* img, int W, int H, int grad_window, float* at,
                         float* dense) {
  Image8 im(W, H);
  std::memcpy(im.data.data(), img, size_t(W) * H);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) at[size_t(y) * W + x] = shi_tomasi_at(im, x, y, grad_window);
  const std::vector<float> d = shi_tomasi_response(im, grad_window);
  std::memcpy(dense, d.data(), d.size() * sizeof(float));
}

// refine_disparity (match.hpp:293-296, match.cpp:768-810) on matches in place.
void ref_refine_disparity(const uint8_t* L, const uint8_t* R, int W, int H,
                          const OrcKeypoint* kl, int nl, const OrcKeypoint* kr, int nr,
                          OrcMatch* m, int n, int half) {
  Image8 li(W, H), ri(W, H);
  std::memcpy(li.data.data(), L, size_t(W) * H);
  std::memcpy(ri.data.data(), R, size_t(W) * H);
  std::vector<Keypoint> a(static_cast<size_t>(nl)), b(static_cast<size_t>(nr));
  for (int i = 0; i < nl; ++i) a[size_t(i)] = Keypoint{kl[i].x, kl[i].y, kl[i].response, kl[i].local_std};
  for (int i = 0; i < nr; ++i) b[size_t(i)] = Keypoint{kr[i].x, kr[i].y, kr[i].response, kr[i].local_std};
  std::vector<Match> v(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    Match& q = v[size_t(i)];
    q.left = m[i].left;
    q.right = m[i].right;
    q.score = m[i].score;
    q.belief = m[i].belief;
    q.disparity = m[i].disparity;
    q.dy = m[i].dy;
    q.margin = m[i].margin;
  }
  refine_disparity(li, ri, a, b, &v, half);
  for (int i = 0; i < n; ++i) m[i].disparity = v[size_t(i)].disparity;
}

// detect_keypoints (preproc.hpp:160-161), the dense detector, for coverage
// comparisons.
int ref_detect_keypoints_dense(const uint8_t* img, int W, int H, const int* icfg,
                               const float* fcfg, OrcKeypoint* out, int cap, int* n_out) {
  Image8 im(W, H);
  std::memcpy(im.data.data(), img, size_t(W) * H);
  DetectorConfig c;
  c.grad_window = icfg[0];
  c.cell = icfg[1];
  c.per_cell = icfg[2];
  c.min_response = fcfg[0];
  c.min_local_std = fcfg[1];
  c.border = icfg[3];
  c.subpixel = icfg[4] != 0;
  c.fast_threshold = icfg[5];
  c.nms_radius = icfg[6];
  const std::vector<Keypoint> k = detect_keypoints(im, c);
  if (int(k.size()) > cap) return -1;
  for (size_t i = 0; i < k.size(); ++i) {
    out[i].x = k[i].x;
    out[i].y = k[i].y;
    out[i].response = k[i].response;
    out[i].local_std = k[i].local_std;
  }
  *n_out = int(k.size());
  return 0;
}

float ref_matching_objective(const OrcMatch* m, int n, const OrcMatchCfg* cfg,
                             int n_left, int n_right) {
  Cfg c;
  to_cfg(cfg, nullptr, nullptr, &c);
  std::vector<Match> v(static_cast<size_t>(n));
  for (int k = 0; k < n; ++k) v[size_t(k)].score = m[k].score;
  return matching_objective(v, c.c, n_left, n_right);
}

}  // extern "C"
