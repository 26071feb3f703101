// CPU restatement of the keypoint-graph BP workload (see knn_bp_ref.h).
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never shipped.
//
// The reference has no k-NN BP solver; its sparse matcher is bipartite
// (core/src/match.cpp). This file is the definition the HIP kernels follow,
// written plainly:
//   labels     for keypoint i at (x, y): d in [0, dmax] with x - d >= hw;
//              h(d) = popcount(censusL(x, y) ^ censusR(x - d, y)); keep the
//              n_cand smallest by (h, d); unary cost c = w_data * h.
//   graph      j is a neighbour of i when j is among the k nearest of i or i
//              among the k nearest of j (squared distance, ties by index).
//   BP         min-sum, all messages updated together from the previous ones:
//                B_i(l)   = c_i(l) + sum_{k in N(i)} m_{k->i}(l)   (ascending k)
//                m_{i->j}(q) = min_l [ (B_i(l) - m_{j->i}(l)) + V(l, q) ]
//                V(l, q)  = lambda * min(|d_i(l) - d_j(q)|, trunc)
//              then m_{i->j} -= min_q m_{i->j}(q). A node without labels sends
//              zeros. Messages start at zero.
//   output     after `iters` updates, B_i; best and second label by (B, l).
#include "knn_bp_ref.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "cpu_ref.h"

namespace {

constexpr int kLab = 8;

}  // namespace

extern "C" {

void orc_kb_default_cfg(OrcKnnCfg* c) {
  c->census_hw = 4;
  c->census_hh = 3;
  c->dmax = 127;
  c->n_cand = 8;
  c->k_nn = 8;
  c->iters = 10;
  c->w_data = 1.0f;
  c->lambda = 1.0f;
  c->trunc = 6;
}

int orc_kb_candidates(const uint8_t* L, const uint8_t* R, int W, int H, const int32_t* kp,
                      int n, const OrcKnnCfg* c, int32_t* cand_d, float* cand_c,
                      int32_t* ncand) {
  std::vector<uint64_t> cl(size_t(W) * H), cr(size_t(W) * H);
  orc_census(L, W, H, c->census_hw, c->census_hh, cl.data());
  orc_census(R, W, H, c->census_hw, c->census_hh, cr.data());
  const int nc = std::min(std::max(c->n_cand, 0), kLab);
  std::vector<std::pair<int, int>> hd;
  for (int i = 0; i < n; ++i) {
    const int x = kp[2 * i], y = kp[2 * i + 1];
    if (x < c->census_hw || x >= W - c->census_hw || y < c->census_hh ||
        y >= H - c->census_hh)
      return -1;
    const uint64_t a = cl[size_t(y) * W + x];
    hd.clear();
    for (int d = 0; d <= c->dmax && x - d >= c->census_hw; ++d)
      hd.emplace_back(__builtin_popcountll(a ^ cr[size_t(y) * W + (x - d)]), d);
    std::sort(hd.begin(), hd.end());
    const int m = std::min(nc, int(hd.size()));
    ncand[i] = m;
    for (int l = 0; l < kLab; ++l) {
      cand_d[i * kLab + l] = l < m ? hd[size_t(l)].second : -1;
      cand_c[i * kLab + l] = l < m ? c->w_data * float(hd[size_t(l)].first) : 0.f;
    }
  }
  return 0;
}

int orc_kb_knn_graph(const int32_t* kp, int n, int k, int32_t* off, int32_t* adj, int cap) {
  std::vector<std::vector<int>> nb(static_cast<size_t>(std::max(n, 0)));
  std::vector<std::pair<int64_t, int>> dist;
  const int kk = std::min(k, std::max(n - 1, 0));
  for (int i = 0; i < n; ++i) {
    dist.clear();
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      const int64_t dx = kp[2 * j] - kp[2 * i], dy = kp[2 * j + 1] - kp[2 * i + 1];
      dist.emplace_back(dx * dx + dy * dy, j);
    }
    std::partial_sort(dist.begin(), dist.begin() + kk, dist.end());
    for (int t = 0; t < kk; ++t) {
      nb[size_t(i)].push_back(dist[size_t(t)].second);
      nb[size_t(dist[size_t(t)].second)].push_back(i);
    }
  }
  int e = 0;
  off[0] = 0;
  for (int i = 0; i < n; ++i) {
    std::vector<int>& v = nb[size_t(i)];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    for (int j : v) {
      if (e >= cap) return -1;
      adj[e++] = j;
    }
    off[i + 1] = e;
  }
  return e;
}

int orc_kb_bp_graph(int n, const int32_t* cand_d, const float* cand_c, const int32_t* ncand,
                    const int32_t* off, const int32_t* adj, const OrcKnnCfg* c,
                    OrcKnnOut* out, float* beliefs) {
  const int E = off[n];
  // rev[e]: the edge j->i of e = i->j
  std::vector<int> rev(size_t(E), -1);
  for (int i = 0; i < n; ++i)
    for (int e = off[i]; e < off[i + 1]; ++e) {
      const int j = adj[e];
      for (int f = off[j]; f < off[j + 1]; ++f)
        if (adj[f] == i) rev[size_t(e)] = f;
      if (rev[size_t(e)] < 0) return -1;   // not symmetric
    }
  std::vector<float> msg(size_t(E) * kLab, 0.f), nxt(size_t(E) * kLab, 0.f);
  std::vector<float> B(kLab);
  auto belief = [&](int i, const std::vector<float>& m, float* b) {
    for (int l = 0; l < ncand[i]; ++l) {
      float v = cand_c[i * kLab + l];
      for (int e = off[i]; e < off[i + 1]; ++e) v = v + m[size_t(rev[size_t(e)]) * kLab + l];
      b[l] = v;
    }
  };
  for (int it = 0; it < c->iters; ++it) {
    for (int i = 0; i < n; ++i) {
      belief(i, msg, B.data());
      const int ni = ncand[i];
      for (int e = off[i]; e < off[i + 1]; ++e) {
        const int j = adj[e], nj = ncand[j];
        float* o = &nxt[size_t(e) * kLab];
        for (int q = 0; q < kLab; ++q) o[q] = 0.f;
        if (ni == 0 || nj == 0) continue;
        const float* back = &msg[size_t(rev[size_t(e)]) * kLab];
        for (int q = 0; q < nj; ++q) {
          float best = std::numeric_limits<float>::infinity();
          for (int l = 0; l < ni; ++l) {
            const int dd = std::abs(cand_d[i * kLab + l] - cand_d[j * kLab + q]);
            const float v = (B[size_t(l)] - back[l]) + c->lambda * float(std::min(dd, c->trunc));
            best = std::min(best, v);
          }
          o[q] = best;
        }
        float mn = o[0];
        for (int q = 1; q < nj; ++q) mn = std::min(mn, o[q]);
        for (int q = 0; q < nj; ++q) o[q] = o[q] - mn;
      }
    }
    msg.swap(nxt);
  }
  const float inf = std::numeric_limits<float>::infinity();
  for (int i = 0; i < n; ++i) {
    belief(i, msg, B.data());
    int l1 = -1, l2 = -1;
    for (int l = 0; l < ncand[i]; ++l) {
      if (l1 < 0 || B[size_t(l)] < B[size_t(l1)]) { l2 = l1; l1 = l; }
      else if (l2 < 0 || B[size_t(l)] < B[size_t(l2)]) l2 = l;
    }
    out[i].d1 = l1 >= 0 ? cand_d[i * kLab + l1] : -1;
    out[i].c1 = l1 >= 0 ? B[size_t(l1)] : inf;
    out[i].d2 = l2 >= 0 ? cand_d[i * kLab + l2] : -1;
    out[i].c2 = l2 >= 0 ? B[size_t(l2)] : inf;
    if (beliefs)
      for (int l = 0; l < kLab; ++l) beliefs[i * kLab + l] = l < ncand[i] ? B[size_t(l)] : inf;
  }
  return 0;
}

int orc_kb_run(const uint8_t* L, const uint8_t* R, int W, int H, const int32_t* kp, int n,
               const OrcKnnCfg* c, OrcKnnOut* out) {
  std::vector<int32_t> cd(size_t(n) * kLab), nc(static_cast<size_t>(n));
  std::vector<float> cc(size_t(n) * kLab);
  if (orc_kb_candidates(L, R, W, H, kp, n, c, cd.data(), cc.data(), nc.data())) return -1;
  const int cap = n * (2 * c->k_nn + 1) + 1;
  std::vector<int32_t> off(size_t(n) + 1), adj(static_cast<size_t>(cap));
  if (orc_kb_knn_graph(kp, n, c->k_nn, off.data(), adj.data(), cap) < 0) return -1;
  return orc_kb_bp_graph(n, cd.data(), cc.data(), nc.data(), off.data(), adj.data(), c, out,
                         nullptr);
}

int orc_kb_brute_min_marginals(int n, const int32_t* cand_d, const float* cand_c,
                               const int32_t* ncand, const int32_t* off, const int32_t* adj,
                               const OrcKnnCfg* c, double* mm) {
  for (int i = 0; i < n * kLab; ++i) mm[i] = std::numeric_limits<double>::infinity();
  std::vector<int> lab(size_t(n), 0);
  for (int i = 0; i < n; ++i)
    if (ncand[i] < 1) return -1;
  for (;;) {
    double e = 0.0;
    for (int i = 0; i < n; ++i) e += cand_c[i * kLab + lab[size_t(i)]];
    for (int i = 0; i < n; ++i)
      for (int f = off[i]; f < off[i + 1]; ++f) {
        const int j = adj[f];
        if (j <= i) continue;
        const int dd = std::abs(cand_d[i * kLab + lab[size_t(i)]] - cand_d[j * kLab + lab[size_t(j)]]);
        e += double(c->lambda) * double(std::min(dd, c->trunc));
      }
    for (int i = 0; i < n; ++i) {
      double& v = mm[i * kLab + lab[size_t(i)]];
      v = std::min(v, e);
    }
    int p = 0;
    while (p < n && ++lab[size_t(p)] == ncand[p]) lab[size_t(p++)] = 0;
    if (p == n) break;
  }
  return 0;
}

}  // extern "C"
