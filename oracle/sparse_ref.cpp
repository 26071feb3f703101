// CPU restatement of the reference's sparse MASDA matcher.
// TEST INFRASTRUCTURE ONLY: the checker for the HIP sparse path, never shipped.
//
// Follows core/src/match.cpp:
//   candidate graph        build_candidates   match.cpp:46-102
//   pair score             pair_score         match.cpp:295-332
//   prior lookup           DisparityPrior     match.cpp:235-244
//   row margins            row_margins        match.cpp:112-132
//   message passing        match_masda        match.cpp:334-417
//   decision               emit               match.cpp:134-231
//   mutual-NN baseline     match_mutual_nn    match.cpp:419-476
//   exhaustive search      match_brute_force  match.cpp:478-543
//   objective              matching_objective match.cpp:812-819
// Built with -ffp-contract=off like the reference, so every float expression
// rounds as the reference's does.
#include "sparse_ref.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>
#include <vector>

namespace {

constexpr float kNeg = -1e30f;
constexpr int kCensusBits = 48;   // match.cpp:340, fixed for the 7x7 descriptor

struct E {
  int i, j;
  float s;
};

int popcount64(uint64_t v) { return __builtin_popcountll(v); }

// max(a, b) as std::max evaluates it: b only when a < b.
float fmax_ref(float a, float b) { return (a < b) ? b : a; }

bool prior_lookup(const OrcPrior* p, float x, float y, float* d, float* sd) {
  if (!p || p->cols <= 0 || p->rows <= 0) return false;
  int cx = int(x) / p->cell;
  int cy = int(y) / p->cell;
  cx = std::min(p->cols - 1, std::max(0, cx));
  cy = std::min(p->rows - 1, std::max(0, cy));
  const size_t k = size_t(cy) * size_t(p->cols) + size_t(cx);
  if (!p->valid[k]) return false;
  *d = p->disparity[k];
  *sd = p->sigma[k];
  return true;
}

// s(i, j): descriptor agreement in [-1, 1] minus the y-residual, prior and
// smoothness penalties (match.cpp:295-332).
float score_of(const OrcKeypoint& l, const OrcKeypoint& r, uint64_t a, uint64_t b,
               const OrcMatchCfg& c, const OrcPrior* prior, const OrcSmooth* sm,
               int li) {
  const float half = 0.5f * float(kCensusBits);
  const float desc = (half - float(popcount64(a ^ b))) / half;
  const float dy = (l.y - r.y) / (c.sigma_y > 0 ? c.sigma_y : 1.f);
  float s = desc - c.w_y * dy * dy;
  float dp = 0.f, sd = 1.f;
  if (prior && prior_lookup(prior, l.x, l.y, &dp, &sd) && sd > 0.f) {
    const float e = ((l.x - r.x) - dp) / sd;
    s -= c.w_prior * e * e;
  }
  if (sm && li >= 0 && li < sm->n && sm->valid[li]) {
    const float sg = sm->sigma[li];
    if (sg > 0.f) {
      const float e = ((l.x - r.x) - sm->disparity[li]) / sg;
      s -= c.w_smooth * e * e;
    }
  }
  return s;
}

// The candidate graph: per left keypoint, right keypoints inside the y band
// and the disparity window, at most max_candidates of them by score.
std::vector<E> candidates(const OrcKeypoint* L, const uint64_t* dl, int nl,
                          const OrcKeypoint* R, const uint64_t* dr, int nr,
                          const OrcMatchCfg& c, const OrcPrior* prior,
                          const OrcSmooth* sm, bool canonical) {
  std::vector<int> by_y(size_t(std::max(nr, 0)));
  std::iota(by_y.begin(), by_y.end(), 0);
  auto y_less = [&](int a, int b) { return R[a].y < R[b].y; };
  if (canonical) std::stable_sort(by_y.begin(), by_y.end(), y_less);
  else std::sort(by_y.begin(), by_y.end(), y_less);

  const size_t cap = size_t(std::max(1, c.max_candidates));
  std::vector<E> out;
  std::vector<std::pair<float, int>> band;
  for (int i = 0; i < nl; ++i) {
    const OrcKeypoint& l = L[i];
    const float ylo = l.y - c.max_dy, yhi = l.y + c.max_dy;
    // first position with y >= ylo, first position with y > yhi
    const size_t b0 = size_t(std::partition_point(by_y.begin(), by_y.end(),
                                                  [&](int j) { return R[j].y < ylo; }) -
                             by_y.begin());
    const size_t b1 = size_t(std::partition_point(by_y.begin(), by_y.end(),
                                                  [&](int j) { return !(yhi < R[j].y); }) -
                             by_y.begin());
    band.clear();
    for (size_t p = b0; p < b1; ++p) {
      const int j = by_y[p];
      const float d = l.x - R[j].x;
      if (d < c.min_disparity || d > c.max_disparity) continue;
      band.emplace_back(score_of(l, R[j], dl[i], dr[j], c, prior, sm, i), j);
    }
    if (band.empty()) continue;
    const size_t keep = std::min(band.size(), cap);
    if (band.size() > keep) {
      auto better = [](const std::pair<float, int>& a, const std::pair<float, int>& b) {
        return a.first > b.first;
      };
      if (canonical) std::stable_sort(band.begin(), band.end(), better);
      else std::partial_sort(band.begin(), band.begin() + long(keep), band.end(), better);
    }
    for (size_t k = 0; k < keep; ++k) out.push_back({i, band[k].second, band[k].first});
  }
  return out;
}

// Best-minus-runner-up score per left keypoint, against lambda when there is
// no runner-up (match.cpp:112-132).
std::vector<float> margins(const std::vector<E>& es, int nl, float lambda) {
  std::vector<float> top(size_t(nl), kNeg), run(size_t(nl), kNeg), m(size_t(nl), 0.f);
  for (const E& e : es) {
    float& t = top[size_t(e.i)];
    float& r = run[size_t(e.i)];
    if (e.s > t) { r = t; t = e.s; }
    else if (e.s > r) r = e.s;
  }
  for (int i = 0; i < nl; ++i) {
    if (top[size_t(i)] == kNeg) continue;
    const float alt = run[size_t(i)] == kNeg ? lambda : fmax_ref(run[size_t(i)], lambda);
    m[size_t(i)] = top[size_t(i)] - alt;
  }
  return m;
}

OrcMatch make_match(const E& e, float belief, const std::vector<float>& mg,
                    const OrcKeypoint* L, const OrcKeypoint* R) {
  OrcMatch m;
  m.left = e.i;
  m.right = e.j;
  m.score = e.s;
  m.belief = belief;
  m.margin = mg[size_t(e.i)];
  m.disparity = L[e.i].x - R[e.j].x;
  m.dy = L[e.i].y - R[e.j].y;
  return m;
}

// Decision (match.cpp:134-231): mutual row/column agreement on the belief,
// then greedy completion in belief order; an edge is eligible only when its
// score beats lambda.
std::vector<OrcMatch> decide(const std::vector<E>& es, const std::vector<float>& bel,
                             int nl, int nr, const OrcMatchCfg& c,
                             const OrcKeypoint* L, const OrcKeypoint* R, bool canonical) {
  std::vector<int> row_pick(size_t(nl), -1), col_pick(size_t(nr), -1);
  for (int e = 0; e < int(es.size()); ++e) {
    if (es[size_t(e)].s <= c.lambda) continue;
    int& rp = row_pick[size_t(es[size_t(e)].i)];
    int& cp = col_pick[size_t(es[size_t(e)].j)];
    if (rp < 0 || bel[size_t(e)] > bel[size_t(rp)]) rp = e;
    if (cp < 0 || bel[size_t(e)] > bel[size_t(cp)]) cp = e;
  }
  const std::vector<float> mg = margins(es, nl, c.lambda);
  std::vector<OrcMatch> out;
  std::vector<char> lt(size_t(nl), 0), rt(size_t(nr), 0);
  auto take = [&](int e) {
    out.push_back(make_match(es[size_t(e)], bel[size_t(e)], mg, L, R));
    lt[size_t(es[size_t(e)].i)] = 1;
    rt[size_t(es[size_t(e)].j)] = 1;
  };
  for (int i = 0; i < nl; ++i) {
    const int e = row_pick[size_t(i)];
    if (e >= 0 && col_pick[size_t(es[size_t(e)].j)] == e) take(e);
  }
  std::vector<int> rest;
  for (int e = 0; e < int(es.size()); ++e) {
    const E& x = es[size_t(e)];
    if (x.s > c.lambda && !lt[size_t(x.i)] && !rt[size_t(x.j)]) rest.push_back(e);
  }
  auto order = [&](int a, int b) {
    if (bel[size_t(a)] != bel[size_t(b)]) return bel[size_t(a)] > bel[size_t(b)];
    return es[size_t(a)].s > es[size_t(b)].s;
  };
  if (canonical) std::stable_sort(rest.begin(), rest.end(), order);
  else std::sort(rest.begin(), rest.end(), order);
  for (int e : rest)
    if (!lt[size_t(es[size_t(e)].i)] && !rt[size_t(es[size_t(e)].j)]) take(e);
  return out;
}

int copy_out(const std::vector<OrcMatch>& v, OrcMatch* out, int* n_out) {
  if (out && !v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(OrcMatch));
  if (n_out) *n_out = int(v.size());
  return 0;
}

void stats_init(OrcMatchStats* st, int nl, int nr, int ne, int converged) {
  if (!st) return;
  st->left = nl;
  st->right = nr;
  st->candidates = ne;
  st->iterations_run = 0;
  st->converged = converged;
  st->final_max_delta = 0.f;
  st->oscillating = 0;
}

}  // namespace

extern "C" {

void orc_match_default_cfg(OrcMatchCfg* c) {
  // match.hpp:136-176 defaults
  c->max_dy = 2.0f;
  c->min_disparity = 1.0f;
  c->max_disparity = 220.0f;
  c->max_candidates = 24;
  c->sigma_y = 1.0f;
  c->w_y = 1.0f;
  c->w_prior = 1.0f;
  c->w_smooth = 1.0f;
  c->iterations = 20;
  c->damping = 0.4f;
  c->lambda = 0.0f;
  c->gamma = 0.0f;
  c->converge_eps = 1e-4f;
}

int orc_match_candidates(const OrcKeypoint* L, const uint64_t* dl, int nl,
                         const OrcKeypoint* R, const uint64_t* dr, int nr,
                         const OrcMatchCfg* cfg, const OrcPrior* prior,
                         const OrcSmooth* sm, int canonical, int* ei, int* ej, float* es) {
  const std::vector<E> g = candidates(L, dl, nl, R, dr, nr, *cfg, prior, sm, canonical != 0);
  for (size_t e = 0; e < g.size(); ++e) {
    ei[e] = g[e].i;
    ej[e] = g[e].j;
    es[e] = g[e].s;
  }
  return int(g.size());
}

int orc_match_masda(const OrcKeypoint* L, const uint64_t* dl, int nl,
                    const OrcKeypoint* R, const uint64_t* dr, int nr,
                    const OrcMatchCfg* cfg, const OrcPrior* prior,
                    const OrcSmooth* sm, int canonical,
                    OrcMatch* out, int* n_out, OrcMatchStats* st) {
  const OrcMatchCfg& c = *cfg;
  const std::vector<E> es = candidates(L, dl, nl, R, dr, nr, c, prior, sm, canonical != 0);
  const int ne = int(es.size());
  stats_init(st, nl, nr, ne, 0);
  if (ne == 0) return copy_out({}, out, n_out);

  // Edge lists per row (left) and per column (right), in edge order.
  std::vector<std::vector<int>> rows(static_cast<size_t>(nl)), cols(static_cast<size_t>(nr));
  for (int e = 0; e < ne; ++e) {
    rows[size_t(es[size_t(e)].i)].push_back(e);
    cols[size_t(es[size_t(e)].j)].push_back(e);
  }
  std::vector<float> rho(size_t(ne), 0.f), beta(size_t(ne), 0.f);
  const float dmp = std::min(std::max(c.damping, 0.f), 0.99f);

  // One damped half-step over a group: every edge's new message is
  // s - max(floor, max over the OTHER edges of the group of `in`).
  auto half_step = [&](const std::vector<std::vector<int>>& groups,
                       const std::vector<float>& in, std::vector<float>& msg,
                       float floor_v, float& delta) {
    for (const std::vector<int>& g : groups) {
      float best = kNeg, second = kNeg;
      int arg = -1;
      for (int e : g) {
        const float v = in[size_t(e)];
        if (v > best) { second = best; best = v; arg = e; }
        else if (v > second) second = v;
      }
      for (int e : g) {
        const float other = (e == arg) ? second : best;
        const float target = es[size_t(e)].s - fmax_ref(floor_v, other);
        const float nv = (1.f - dmp) * target + dmp * msg[size_t(e)];
        delta = fmax_ref(delta, std::fabs(nv - msg[size_t(e)]));
        msg[size_t(e)] = nv;
      }
    }
  };

  float last = 1e30f;
  for (int it = 0; it < c.iterations; ++it) {
    float delta = 0.f;
    half_step(rows, beta, rho, c.lambda, delta);   // rho: competition within a row
    half_step(cols, rho, beta, c.gamma, delta);    // beta: competition within a column
    if (st) {
      if (delta > last) ++st->oscillating;
      st->iterations_run = it + 1;
      st->final_max_delta = delta;
    }
    last = delta;
    if (delta < c.converge_eps) {
      if (st) st->converged = 1;
      break;
    }
  }
  std::vector<float> bel(static_cast<size_t>(ne));
  for (int e = 0; e < ne; ++e)
    bel[size_t(e)] = beta[size_t(e)] + rho[size_t(e)] - es[size_t(e)].s;
  return copy_out(decide(es, bel, nl, nr, c, L, R, canonical != 0), out, n_out);
}

int orc_match_mutual_nn(const OrcKeypoint* L, const uint64_t* dl, int nl,
                        const OrcKeypoint* R, const uint64_t* dr, int nr,
                        const OrcMatchCfg* cfg, const OrcPrior* prior,
                        const OrcSmooth* sm, float ratio, int canonical,
                        OrcMatch* out, int* n_out, OrcMatchStats* st) {
  const OrcMatchCfg& c = *cfg;
  const std::vector<E> es = candidates(L, dl, nl, R, dr, nr, c, prior, sm, canonical != 0);
  const int ne = int(es.size());
  stats_init(st, nl, nr, ne, 1);
  if (ne == 0) return copy_out({}, out, n_out);
  const std::vector<float> mg = margins(es, nl, c.lambda);
  std::vector<int> b1(size_t(nl), -1), b2(size_t(nl), -1), bc(size_t(nr), -1);
  for (int e = 0; e < ne; ++e) {
    const E& x = es[size_t(e)];
    int& r1 = b1[size_t(x.i)];
    int& r2 = b2[size_t(x.i)];
    if (r1 < 0 || x.s > es[size_t(r1)].s) { r2 = r1; r1 = e; }
    else if (r2 < 0 || x.s > es[size_t(r2)].s) r2 = e;
    int& cb = bc[size_t(x.j)];
    if (cb < 0 || x.s > es[size_t(cb)].s) cb = e;
  }
  std::vector<OrcMatch> v;
  for (int i = 0; i < nl; ++i) {
    const int e = b1[size_t(i)];
    if (e < 0 || es[size_t(e)].s <= c.lambda || bc[size_t(es[size_t(e)].j)] != e) continue;
    if (b2[size_t(i)] >= 0) {
      const float b = es[size_t(e)].s - c.lambda;
      const float s2 = es[size_t(b2[size_t(i)])].s - c.lambda;
      if (b > 0.f && s2 > 0.f && s2 / b > ratio) continue;
    }
    v.push_back(make_match(es[size_t(e)], es[size_t(e)].s, mg, L, R));
  }
  return copy_out(v, out, n_out);
}

int orc_match_brute_force(const OrcKeypoint* L, const uint64_t* dl, int nl,
                          const OrcKeypoint* R, const uint64_t* dr, int nr,
                          const OrcMatchCfg* cfg, OrcMatch* out, int* n_out) {
  const OrcMatchCfg& c = *cfg;
  const std::vector<E> es = candidates(L, dl, nl, R, dr, nr, c, nullptr, nullptr, false);
  std::vector<float> s(size_t(nl) * size_t(nr), 0.f);
  std::vector<char> has(size_t(nl) * size_t(nr), 0);
  for (const E& e : es) {
    s[size_t(e.i) * size_t(nr) + size_t(e.j)] = e.s;
    has[size_t(e.i) * size_t(nr) + size_t(e.j)] = 1;
  }
  const std::vector<float> mg = margins(es, nl, c.lambda);
  std::vector<int> cur(size_t(nl), -1), best(size_t(nl), -1);
  std::vector<char> used(size_t(nr), 0);
  float best_v = 0.f;
  for (int i = 0; i < nl; ++i) best_v += c.lambda;
  // Depth-first over left keypoints, in the reference's float summation order.
  std::function<void(int, float)> dfs = [&](int i, float acc) {
    if (i == nl) {
      float tot = acc;
      for (int j = 0; j < nr; ++j)
        if (!used[size_t(j)]) tot += c.gamma;
      if (tot > best_v) { best_v = tot; best = cur; }
      return;
    }
    cur[size_t(i)] = -1;
    dfs(i + 1, acc + c.lambda);
    for (int j = 0; j < nr; ++j) {
      if (used[size_t(j)] || !has[size_t(i) * size_t(nr) + size_t(j)]) continue;
      used[size_t(j)] = 1;
      cur[size_t(i)] = j;
      dfs(i + 1, acc + s[size_t(i) * size_t(nr) + size_t(j)]);
      cur[size_t(i)] = -1;
      used[size_t(j)] = 0;
    }
  };
  dfs(0, 0.f);
  std::vector<OrcMatch> v;
  for (int i = 0; i < nl; ++i) {
    if (best[size_t(i)] < 0) continue;
    const E e{i, best[size_t(i)], s[size_t(i) * size_t(nr) + size_t(best[size_t(i)])]};
    v.push_back(make_match(e, e.s, mg, L, R));
  }
  return copy_out(v, out, n_out);
}

float orc_matching_objective(const OrcMatch* m, int n, const OrcMatchCfg* cfg,
                             int n_left, int n_right) {
  float t = 0.f;
  for (int k = 0; k < n; ++k) t += m[k].score;
  t += cfg->lambda * float(n_left - n);
  t += cfg->gamma * float(n_right - n);
  return t;
}

}  // extern "C"
