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

float ref_matching_objective(const OrcMatch* m, int n, const OrcMatchCfg* cfg,
                             int n_left, int n_right) {
  Cfg c;
  to_cfg(cfg, nullptr, nullptr, &c);
  std::vector<Match> v(static_cast<size_t>(n));
  for (int k = 0; k < n; ++k) v[size_t(k)].score = m[k].score;
  return matching_objective(v, c.c, n_left, n_right);
}

}  // extern "C"
