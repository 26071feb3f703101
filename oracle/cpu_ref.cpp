// CPU oracle for the DoubleEye dense MASDA hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is a single-threaded restatement of the
// reference algorithm, used by tests/, by __graft_entry__.smoke() and by
// bench.py's cpu_baseline leg as the checker. The product path (the HIP library
// under doubleeye_amd/) never links, loads or calls it.
//
// Every function cites the reference code it restates (paths relative to the
// reference repository root). Semantics follow `de_dense --threads 1`, the only
// deterministic configuration of the reference (one thread walks every disparity
// plane in ascending order).
//
// Pinning (how each part is checked against something other than itself):
//   census          test_preproc invariants + bit-exact end-to-end vs de_dense
//   rf table/filter end-to-end bit-exact vs de_dense --out/--out-conf
//   score tables    integer closed forms, end-to-end vs de_dense
//   top-2 + nbrs    end-to-end vs de_dense --out-conf (s1,s2,d1,d2 exact) and
//                   the sub-pixel disparities in --out
//   row solver      end-to-end vs de_dense --out (bit-exact on the committed
//                   fixtures), scene coverage known answers
//   confidence      test_confidence.cpp golden values
//   fp32 volume     de_dense --dump-vol (float path), within 1e-5 relative
//
// Build: g++ -O2 -std=c++14 -fPIC -shared (see oracle/Makefile). No -march and
// no FMA contraction, like the reference build, so float arithmetic matches.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "cpu_ref.h"

namespace {

const int kQ = 14;
const int kOne = 1 << kQ;              // Q14 unit score
const float kNone = -1e30f;            // "no candidate" sentinel

// census bit b in scan order dy=-hh..hh (outer), dx=-hw..hw (inner), centre
// skipped; set iff neighbour < centre; 0 where the window is cut.
// Follows core/src/preproc.cpp:110-138 and core/tools/de_dense.cpp:94-129.
void census_plane(const uint8_t* img, int W, int H, int hw, int hh, uint64_t* out) {
  std::fill(out, out + size_t(W) * H, 0ull);
  for (int y = hh; y < H - hh; ++y)
    for (int x = hw; x < W - hw; ++x) {
      const uint8_t c = img[size_t(y) * W + x];
      uint64_t v = 0;
      int b = 0;
      for (int dy = -hh; dy <= hh; ++dy)
        for (int dx = -hw; dx <= hw; ++dx) {
          if (!dx && !dy) continue;
          if (img[size_t(y + dy) * W + x + dx] < c) v |= 1ull << b;
          ++b;
        }
      out[size_t(y) * W + x] = v;
    }
}

// Q15 coefficient table, core/tools/de_dense.cpp:531-546.
void rf_table(float sigma_s, float sigma_r, uint16_t* tbl) {
  const float k = -std::sqrt(2.f) / sigma_s;
  const float rs = sigma_s / sigma_r;
  for (int d = 0; d < 256; ++d) {
    const float c = std::exp(k * (1.f + rs * (float(d) / 255.f)));
    tbl[d] = uint16_t(std::min(32767.f, c * 32768.f + 0.5f));
  }
}

// ax: step into x from x-1 (tbl[0] at x=0); ay: step into y from y-1 (tbl[0]
// at y=0). Left image only. core/tools/de_dense.cpp:547-560.
void rf_coeff_planes(const uint8_t* L, int W, int H, const uint16_t* tbl,
                     uint16_t* ax, uint16_t* ay) {
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t i = size_t(y) * W + x;
      ax[i] = x ? tbl[std::abs(int(L[i]) - int(L[i - 1]))] : tbl[0];
      ay[i] = y ? tbl[std::abs(int(L[i]) - int(L[i - W]))] : tbl[0];
    }
}

// One recursive step, int32 carry, round-to-nearest >>15.
// core/tools/de_dense.cpp:411-430.
inline int32_t rf_step(int32_t prev, int32_t c, int32_t a) {
  return c + ((a * (prev - c) + (1 << 14)) >> 15);
}

// Four passes over a whole int16 plane: H-fwd, H-bwd, V-fwd, V-bwd.
// core/tools/de_dense.cpp:485-508 (whole plane, border zeros propagate).
void rf_filter_plane_i16(int16_t* C, const uint16_t* ax, const uint16_t* ay,
                         int W, int H) {
  for (int y = 0; y < H; ++y) {
    int16_t* r = C + size_t(y) * W;
    const uint16_t* a = ax + size_t(y) * W;
    int32_t f = r[0];
    for (int x = 1; x < W; ++x) { f = rf_step(f, r[x], a[x]); r[x] = int16_t(f); }
    f = r[W - 1];
    for (int x = W - 2; x >= 0; --x) { f = rf_step(f, r[x], a[x + 1]); r[x] = int16_t(f); }
  }
  for (int y = 1; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t i = size_t(y) * W + x;
      C[i] = int16_t(rf_step(C[i - W], C[i], ay[i]));
    }
  for (int y = H - 2; y >= 0; --y)
    for (int x = 0; x < W; ++x) {
      const size_t i = size_t(y) * W + x;
      C[i] = int16_t(rf_step(C[i + W], C[i], ay[i + W]));
    }
}

// Same four passes in float, core/tools/de_dense.cpp:379-396, 510-527.
void rf_filter_plane_f32(float* C, const uint16_t* ax, const uint16_t* ay,
                         int W, int H) {
  const float s = 1.f / 32768.f;
  for (int y = 0; y < H; ++y) {
    float* r = C + size_t(y) * W;
    const uint16_t* a = ax + size_t(y) * W;
    float f = r[0];
    for (int x = 1; x < W; ++x) { f = r[x] + float(a[x]) * s * (f - r[x]); r[x] = f; }
    f = r[W - 1];
    for (int x = W - 2; x >= 0; --x) { f = r[x] + float(a[x + 1]) * s * (f - r[x]); r[x] = f; }
  }
  for (int y = 1; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t i = size_t(y) * W + x;
      C[i] += float(ay[i]) * s * (C[i - W] - C[i]);
    }
  for (int y = H - 2; y >= 0; --y)
    for (int x = 0; x < W; ++x) {
      const size_t i = size_t(y) * W + x;
      C[i] += float(ay[i + W]) * s * (C[i + W] - C[i]);
    }
}

// Graded score tables, core/tools/de_dense.cpp:1201-1223: census level table
// ((half-h)*ONE)/half with C++ truncating division, truncated-AD table, and the
// blend weight wq = int(clamp(ad,0,1)*1024).
struct ScoreTables {
  int16_t tbl[65];
  int16_t adt[256];
  int32_t wq;
};

ScoreTables make_score_tables(int half, float ad, int ad_trunc) {
  ScoreTables t;
  std::memset(&t, 0, sizeof(t));
  for (int h = 0; h <= 2 * half && h < 65; ++h)
    t.tbl[h] = int16_t(((half - h) * kOne) / half);
  t.wq = int32_t(std::max(0.f, std::min(1.f, ad)) * 1024.f);
  const int T = std::max(1, ad_trunc);
  for (int v = 0; v < 256; ++v)
    t.adt[v] = int16_t(kOne - (2 * kOne * std::min(v, T)) / T);
  return t;
}

inline int16_t graded(const ScoreTables& t, uint64_t cl, uint64_t cr, uint8_t l,
                      uint8_t r) {
  const int32_t c = t.tbl[__builtin_popcountll(cl ^ cr)];
  if (!t.wq) return int16_t(c);
  const int32_t a = t.adt[std::abs(int(l) - int(r))];
  return int16_t((c * (1024 - t.wq) + a * t.wq) >> 10);
}

struct Top2f {
  float b1 = kNone, b2 = kNone;
  int i1 = -1;
  void push(float v, int i) {
    if (v > b1) { b2 = b1; b1 = v; i1 = i; }
    else if (v > b2) { b2 = v; }
  }
  float excl(int i) const { return i == i1 ? b2 : b1; }
};

// Row solver over K=2 pre-selected candidates: margin, counting-sort buckets by
// right pixel, reverse match, `iters` rounds of rho/beta, belief, greedy claim
// in std::sort order, gates, sub-pixel fit, reverse-match residual.
// Follows core/include/doubleeye/dense_solve.hpp:213-452 (vol == nullptr path).
void solve_row_k2(int W, const OrcCfg& cfg, const float* cs, const int* cd,
                  const int* cn, const float* cnb, float* out_disp,
                  float* out_margin, float* out_lrc, int* tie_flag) {
  const int K = 2;
  const int dmin = cfg.dmin;
  const float keep = 1.f - cfg.damping;
  std::vector<int> bstart(W + 1), bitems(size_t(W) * K), cursor(W);
  std::vector<float> beta(size_t(W) * K, 0.f), rho(size_t(W) * K, 0.f);
  std::vector<float> best(W);
  std::vector<int> bestj(W), order(W);
  std::vector<char> taken(W);

  for (int x = 0; x < W; ++x) {
    const float* S = cs + size_t(x) * K;
    const int n = cn[x];
    const float alt = (n < 2) ? cfg.lambda : std::max(S[1], cfg.lambda);
    out_margin[x] = (n == 0) ? 0.f : S[0] - alt;
  }

  std::fill(cursor.begin(), cursor.end(), 0);
  for (int x = 0; x < W; ++x)
    for (int j = 0; j < cn[x]; ++j) {
      const int xr = x - (dmin + cd[size_t(x) * K + j]);
      if (xr >= 0 && xr < W) ++cursor[xr];
    }
  bstart[0] = 0;
  for (int xr = 0; xr < W; ++xr) bstart[xr + 1] = bstart[xr] + cursor[xr];
  for (int xr = 0; xr < W; ++xr) cursor[xr] = bstart[xr];
  for (int x = 0; x < W; ++x)
    for (int j = 0; j < cn[x]; ++j) {
      const int xr = x - (dmin + cd[size_t(x) * K + j]);
      if (xr >= 0 && xr < W) bitems[cursor[xr]++] = x * K + j;
    }

  // reverse match: best claimant (first of equals) per right pixel
  std::vector<int> rev(W, -1);
  for (int xr = 0; xr < W; ++xr) {
    float bv = kNone;
    for (int p = bstart[xr]; p < bstart[xr + 1]; ++p) {
      const int idx = bitems[p];
      if (cs[idx] > bv) { bv = cs[idx]; rev[xr] = cd[idx]; }
    }
  }

  for (int it = 0; it < cfg.iters; ++it) {
    for (int x = 0; x < W; ++x) {
      const int n = cn[x];
      if (!n) continue;
      const int b = x * K;
      Top2f t;
      for (int j = 0; j < n; ++j) t.push(beta[b + j], j);
      for (int j = 0; j < n; ++j) {
        const float tgt = cs[b + j] - std::max(cfg.lambda, t.excl(j));
        rho[b + j] = keep * tgt + cfg.damping * rho[b + j];
      }
    }
    for (int xr = 0; xr < W; ++xr) {
      const int lo = bstart[xr], hi = bstart[xr + 1];
      if (hi <= lo) continue;
      Top2f t;
      for (int p = lo; p < hi; ++p) t.push(rho[bitems[p]], p - lo);
      for (int p = lo; p < hi; ++p) {
        const int idx = bitems[p];
        beta[idx] = keep * (cs[idx] - std::max(cfg.gamma, t.excl(p - lo)))
                    + cfg.damping * beta[idx];
      }
    }
  }

  for (int x = 0; x < W; ++x) {
    best[x] = kNone; bestj[x] = -1; taken[x] = 0; order[x] = x;
    const int b = x * K;
    for (int j = 0; j < cn[x]; ++j) {
      if (cs[b + j] <= cfg.lambda) continue;
      const float bel = (cfg.iters > 0) ? (beta[b + j] + rho[b + j] - cs[b + j])
                                        : cs[b + j];
      if (bel > best[x]) { best[x] = bel; bestj[x] = j; }
    }
  }

  // Pixels whose claim competes, at an exactly equal belief, with another
  // claimant of the same right pixel: the greedy winner between them is decided
  // by std::sort's (unstable) order and is not a property of the algorithm.
  if (tie_flag) {
    std::vector<float> top(W, kNone);
    std::vector<int> cnt(W, 0);
    for (int x = 0; x < W; ++x) {
      tie_flag[x] = 0;
      if (bestj[x] < 0) continue;
      const int xr = x - (dmin + cd[size_t(x) * K + bestj[x]]);
      if (xr < 0) continue;
      if (best[x] > top[xr]) { top[xr] = best[x]; cnt[xr] = 1; }
      else if (best[x] == top[xr]) ++cnt[xr];
    }
    for (int x = 0; x < W; ++x) {
      if (bestj[x] < 0) continue;
      const int xr = x - (dmin + cd[size_t(x) * K + bestj[x]]);
      if (xr >= 0 && cnt[xr] > 1 && best[x] == top[xr]) tie_flag[x] = 1;
    }
  }

  std::sort(order.begin(), order.end(),
            [&](int a, int b) { return best[a] > best[b]; });
  for (int oi = 0; oi < W; ++oi) {
    const int x = order[oi];
    out_disp[x] = std::nanf("");
    if (bestj[x] < 0) continue;
    if (cfg.min_margin > 0.f && out_margin[x] < cfg.min_margin) continue;
    if (cfg.min_ratio > 0.f) {
      const int b = x * K;
      const float s2 = (cn[x] > 1) ? std::max(cs[b + 1], cfg.lambda) : cfg.lambda;
      const float c1 = std::max(1.f - cs[b], 1e-6f);
      if ((1.f - s2) / c1 < cfg.min_ratio) continue;
    }
    const int d = dmin + cd[size_t(x) * K + bestj[x]];
    const int xr = x - d;
    if (xr < 0 || taken[xr]) continue;
    taken[xr] = 1;
    float off = 0.f;
    if (cfg.subpixel && cnb && bestj[x] == 0) {
      const float c0 = cs[size_t(x) * K];
      const float cm = cnb[size_t(x) * 2], cp = cnb[size_t(x) * 2 + 1];
      if (cm > -1e29f && cp > -1e29f) {
        if (cfg.fit_eq) {
          const float lower = cm < cp ? cm : cp;
          const float den = c0 - lower;
          if (den > 1e-6f) off = 0.5f * (cp - cm) / den;
        } else {
          const float den = 2.f * c0 - cm - cp;
          if (den > 1e-6f) off = 0.5f * (cp - cm) / den;
        }
        if (off > 0.5f) off = 0.5f;
        if (off < -0.5f) off = -0.5f;
      }
    }
    out_disp[x] = float(d) + off;
  }

  if (out_lrc) {
    for (int x = 0; x < W; ++x) {
      out_lrc[x] = std::nanf("");
      const float dl = out_disp[x];
      if (!(dl == dl)) continue;
      const int xr = x - int(std::lround(dl));
      if (xr >= 0 && xr < W && rev[xr] >= 0)
        out_lrc[x] = std::fabs(dl - float(dmin + rev[xr]));
    }
  }
}

// Candidates -> solve_row_sparse per row, confidence planes, LRC plane
// (de_dense.cpp:1498-1597): the part after the insert, shared by the full and
// the masked sweeps.
void finish_dense(int W, int H, const OrcCfg& cfg, const std::vector<int16_t>& s0,
                  const std::vector<int16_t>& s1, const std::vector<int16_t>& d0,
                  const std::vector<int16_t>& d1, const std::vector<int16_t>& m0,
                  const std::vector<int16_t>& p0, const std::vector<int16_t>& rd, float* disp,
                  float* margin, float* cand6, uint8_t* tie_mask) {
  const size_t WH = size_t(W) * H;
  const float q = 1.f / float(kOne);
  std::vector<float> cs(size_t(W) * 2), cnb(size_t(W) * 2), lrc2(W);
  std::vector<int> cd(size_t(W) * 2), cn(W), tie(W);
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      const size_t i = size_t(y) * W + x;
      int n = 0;
      cs[2 * x] = cs[2 * x + 1] = kNone;
      cd[2 * x] = cd[2 * x + 1] = -1;
      cnb[2 * x] = cnb[2 * x + 1] = kNone;
      if (s0[i] > -32768) {
        cs[2 * x] = float(s0[i]) * q; cd[2 * x] = d0[i]; n = 1;
        if (cfg.subpixel) {
          cnb[2 * x] = (m0[i] == -32768) ? kNone : float(m0[i]) * q;
          cnb[2 * x + 1] = (p0[i] == -32768) ? kNone : float(p0[i]) * q;
        }
        if (s1[i] > -32768) { cs[2 * x + 1] = float(s1[i]) * q; cd[2 * x + 1] = d1[i]; n = 2; }
      }
      cn[x] = n;
    }
    solve_row_k2(W, cfg, cs.data(), cd.data(), cn.data(),
                 cfg.subpixel ? cnb.data() : nullptr, disp + size_t(y) * W,
                 margin + size_t(y) * W, cand6 ? lrc2.data() : nullptr,
                 tie_mask ? tie.data() : nullptr);
    if (tie_mask)
      for (int x = 0; x < W; ++x) tie_mask[size_t(y) * W + x] = uint8_t(tie[x]);
    if (cand6) {
      for (int x = 0; x < W; ++x) {
        const size_t i = size_t(y) * W + x;
        cand6[i] = cn[x] > 0 ? cs[2 * x] : kNone;
        cand6[WH + i] = cn[x] > 1 ? cs[2 * x + 1] : kNone;
        cand6[2 * WH + i] = cn[x] > 0 ? float(cfg.dmin + cd[2 * x]) : kNone;
        cand6[3 * WH + i] = cn[x] > 1 ? float(cfg.dmin + cd[2 * x + 1]) : kNone;
        cand6[5 * WH + i] = lrc2[x];
      }
    }
  }
  if (cand6) {
    // Plane 4 keeps the -1e30 fill unless the reverse scan ran, then NaN where
    // a pixel has no answer or its right pixel no claimant (de_dense.cpp:1581-1597).
    for (size_t i = 0; i < WH; ++i) cand6[4 * WH + i] = cfg.lrc ? std::nanf("") : kNone;
    if (cfg.lrc)
      for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
          const size_t i = size_t(y) * W + x;
          const float dl = disp[i];
          if (!std::isfinite(dl)) continue;
          const int xr = x - int(std::lround(dl));
          if (xr >= 0 && xr < W && rd[size_t(y) * W + xr] >= 0)
            cand6[4 * WH + i] = std::fabs(dl - float(cfg.dmin + rd[size_t(y) * W + xr]));
        }
  }
}

}  // namespace

extern "C" {

void orc_default_cfg(OrcCfg* c) {
  c->dmin = 1; c->dmax = 60; c->iters = 0;
  c->census_hw = 3; c->census_hh = 3;
  c->sigma_s = 8.f; c->sigma_r = 0.2f;
  c->ad = 0.15f; c->ad_trunc = 10;
  c->lambda = -0.1f; c->gamma = -0.1f; c->damping = 0.4f;
  c->min_margin = 0.f; c->min_ratio = 0.f;
  c->subpixel = 1; c->fit_eq = 1; c->lrc = 0;
}

int orc_census(const uint8_t* img, int W, int H, int hw, int hh, uint64_t* out) {
  if (!img || !out || W <= 0 || H <= 0 || hw < 0 || hh < 0) return -1;
  if ((2 * hw + 1) * (2 * hh + 1) - 1 > 64) return -2;
  census_plane(img, W, H, hw, hh, out);
  return 0;
}

void orc_rf_table(float sigma_s, float sigma_r, uint16_t* tbl) {
  rf_table(sigma_s, sigma_r, tbl);
}

void orc_rf_coeffs(const uint8_t* L, int W, int H, float sigma_s, float sigma_r,
                   uint16_t* ax, uint16_t* ay) {
  uint16_t t[256];
  rf_table(sigma_s, sigma_r, t);
  rf_coeff_planes(L, W, H, t, ax, ay);
}

void orc_score_tables(int half, float ad, int ad_trunc, int16_t* tbl65,
                      int16_t* adt256, int32_t* wq) {
  const ScoreTables t = make_score_tables(half, ad, ad_trunc);
  std::memcpy(tbl65, t.tbl, sizeof(t.tbl));
  std::memcpy(adt256, t.adt, sizeof(t.adt));
  *wq = t.wq;
}

// Filtered int16 (Q14) planes for disparities [k0, k1), layout [k][y][x].
// The intermediate the device volume must equal bit for bit.
int orc_filtered_planes_i16(const uint8_t* L, const uint8_t* R, int W, int H,
                            const OrcCfg* cfg, int k0, int k1, int16_t* out) {
  const int hw = cfg->census_hw, hh = cfg->census_hh;
  const int nb = (2 * hw + 1) * (2 * hh + 1) - 1;
  if (nb > 64 || (nb & 1)) return -2;
  const size_t WH = size_t(W) * H;
  std::vector<uint64_t> cl(WH), cr(WH);
  census_plane(L, W, H, hw, hh, cl.data());
  census_plane(R, W, H, hw, hh, cr.data());
  std::vector<uint16_t> ax(WH), ay(WH);
  orc_rf_coeffs(L, W, H, cfg->sigma_s, cfg->sigma_r, ax.data(), ay.data());
  const ScoreTables st = make_score_tables(nb / 2, cfg->ad, cfg->ad_trunc);
  for (int k = k0; k < k1; ++k) {
    int16_t* P = out + size_t(k - k0) * WH;
    std::fill(P, P + WH, int16_t(0));
    const int d = cfg->dmin + k;
    for (int y = hh; y < H - hh; ++y)
      for (int x = hw + d; x < W - hw; ++x) {
        const size_t i = size_t(y) * W + x;
        P[i] = graded(st, cl[i], cr[i - d], L[i], R[i - d]);
      }
    rf_filter_plane_i16(P, ax.data(), ay.data(), W, H);
  }
  return 0;
}

// The default dense pipeline of `de_dense --threads 1` (recursive filter,
// blockwise top-2, sub-pixel from plane neighbours), core/tools/de_dense.cpp:
// census 747-748, coefficients 916-923, cost stage 1010-1300 (scalar insert path
// with sub-pixel capture at 1249-1294), per-row 2-of-2n merge 1498-1538 (the
// identity at one thread), solve 1539-1560, reverse-match plane 1581-1597.
//
// Outputs: disp W*H (NaN = no answer), margin W*H, and when cand6 is non-null
// six planes s1,s2,d1,d2,lrc,lrc2 (-1e30 = missing candidate; lrc NaN unless
// cfg.lrc). tie_mask (nullable, W*H bytes) flags exact-tie greedy pixels.
int orc_dense_run(const uint8_t* L, const uint8_t* R, int W, int H,
                  const OrcCfg* cfgp, float* disp, float* margin, float* cand6,
                  uint8_t* tie_mask) {
  if (!L || !R || !cfgp || !disp || !margin || W <= 0 || H <= 0) return -1;
  const OrcCfg& cfg = *cfgp;
  const int hw = cfg.census_hw, hh = cfg.census_hh;
  const int nb = (2 * hw + 1) * (2 * hh + 1) - 1;
  if (nb > 64 || (nb & 1)) return -2;
  const int D = cfg.dmax - cfg.dmin + 1;
  if (D < 1 || D > 256) return -3;
  const size_t WH = size_t(W) * H;

  std::vector<uint64_t> cl(WH), cr(WH);
  census_plane(L, W, H, hw, hh, cl.data());
  census_plane(R, W, H, hw, hh, cr.data());
  std::vector<uint16_t> ax(WH), ay(WH);
  orc_rf_coeffs(L, W, H, cfg.sigma_s, cfg.sigma_r, ax.data(), ay.data());
  const ScoreTables st = make_score_tables(nb / 2, cfg.ad, cfg.ad_trunc);

  std::vector<int16_t> s0(WH, -32768), s1(WH, -32768), d0(WH, -1), d1(WH, -1);
  std::vector<int16_t> m0(WH, -32768), p0(WH, -32768);
  std::vector<int16_t> rs(cfg.lrc ? WH : 0, -32768), rd(cfg.lrc ? WH : 0, -1);
  std::vector<int16_t> plane[2] = {std::vector<int16_t>(WH), std::vector<int16_t>(WH)};
  for (int k = 0; k < D; ++k) {
    const int d = cfg.dmin + k;
    int16_t* P = plane[k & 1].data();
    const int16_t* prv = plane[(k & 1) ^ 1].data();
    std::fill(P, P + WH, int16_t(0));
    for (int y = hh; y < H - hh; ++y)
      for (int x = hw + d; x < W - hw; ++x) {
        const size_t i = size_t(y) * W + x;
        P[i] = graded(st, cl[i], cr[i - d], L[i], R[i - d]);
      }
    rf_filter_plane_i16(P, ax.data(), ay.data(), W, H);
    const int16_t kk = int16_t(k);
    for (int y = hh; y < H - hh; ++y)
      for (int x = hw + d; x < W - hw; ++x) {
        const size_t i = size_t(y) * W + x;
        const int16_t v = P[i];
        if (cfg.lrc) {
          const size_t j = i - size_t(d);
          if (v > rs[j]) { rs[j] = v; rd[j] = kk; }
        }
        if (cfg.subpixel) {
          if (d0[i] == kk - 1) p0[i] = v;
          if (v > s0[i]) { m0[i] = k > 0 ? prv[i] : int16_t(-32768); p0[i] = -32768; }
        }
        if (v <= s1[i]) continue;
        if (v > s0[i]) { s1[i] = s0[i]; d1[i] = d0[i]; s0[i] = v; d0[i] = kk; }
        else { s1[i] = v; d1[i] = kk; }
      }
  }

  finish_dense(W, H, cfg, s0, s1, d0, d1, m0, p0, rd, disp, margin, cand6, tie_mask);
  return 0;
}

// --prior: the prior as a mask on absolute-disparity planes
// (core/tools/de_dense.cpp:797-853, 1156-1296). Plane k is scored and
// filtered on its rectangle only (the bounding box of the pixels whose band
// |d - round(prior)| <= band admits it, padded by 8 and clipped to the score
// region), the filter starting from the rectangle's own edge values; candidates
// are inserted strictly inside each row's interval; no sub-pixel neighbours.
// `prior` is W*H finite floats.
int orc_dense_run_prior(const uint8_t* L, const uint8_t* R, int W, int H, const OrcCfg* cfgp,
                        const float* prior, int band, float* disp, float* margin,
                        float* cand6, uint8_t* tie_mask) {
  if (!L || !R || !cfgp || !prior || !disp || !margin || W <= 0 || H <= 0) return -1;
  OrcCfg cfg = *cfgp;
  cfg.subpixel = 0;
  const int hw = cfg.census_hw, hh = cfg.census_hh;
  const int nb = (2 * hw + 1) * (2 * hh + 1) - 1;
  if (nb > 64 || (nb & 1)) return -2;
  const int D = cfg.dmax - cfg.dmin + 1;
  if (D < 1 || D > 256) return -3;
  const size_t WH = size_t(W) * H;
  for (size_t i = 0; i < WH; ++i)
    if (!std::isfinite(prior[i])) return -4;

  std::vector<int16_t> mx0(size_t(H) * D, int16_t(W)), mx1(size_t(H) * D, int16_t(-1));
  std::vector<int> pylo(D, H), pyhi(D, -1), pxlo(D, W), pxhi(D, -1);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const int dc = int(std::lround(prior[size_t(y) * W + x]));
      const int klo = std::max(cfg.dmin, dc - band) - cfg.dmin;
      const int khi = std::min(cfg.dmax, dc + band) - cfg.dmin;
      for (int k = klo; k <= khi; ++k) {
        int16_t& a = mx0[size_t(y) * D + k];
        int16_t& b = mx1[size_t(y) * D + k];
        if (x < a) a = int16_t(x);
        if (x > b) b = int16_t(x);
      }
    }
  for (int y = 0; y < H; ++y)
    for (int k = 0; k < D; ++k) {
      const int a = mx0[size_t(y) * D + k], b = mx1[size_t(y) * D + k];
      if (b < a) continue;
      pylo[k] = std::min(pylo[k], y);
      pyhi[k] = std::max(pyhi[k], y);
      pxlo[k] = std::min(pxlo[k], a);
      pxhi[k] = std::max(pxhi[k], b);
    }

  std::vector<uint64_t> cl(WH), cr(WH);
  census_plane(L, W, H, hw, hh, cl.data());
  census_plane(R, W, H, hw, hh, cr.data());
  std::vector<uint16_t> ax(WH), ay(WH);
  orc_rf_coeffs(L, W, H, cfg.sigma_s, cfg.sigma_r, ax.data(), ay.data());
  const ScoreTables st = make_score_tables(nb / 2, cfg.ad, cfg.ad_trunc);

  std::vector<int16_t> s0(WH, -32768), s1(WH, -32768), d0(WH, -1), d1(WH, -1);
  std::vector<int16_t> m0(WH, -32768), p0(WH, -32768);
  std::vector<int16_t> rs(cfg.lrc ? WH : 0, -32768), rd(cfg.lrc ? WH : 0, -1);
  std::vector<int16_t> P(WH);
  const int MPAD = 8;
  for (int k = 0; k < D; ++k) {
    const int d = cfg.dmin + k;
    if (pyhi[k] < pylo[k]) continue;
    const int y0 = std::max(hh, pylo[k] - MPAD), y1 = std::min(H - hh, pyhi[k] + 1 + MPAD);
    const int x0 = std::max(hw + d, pxlo[k] - MPAD), x1 = std::min(W - hw, pxhi[k] + 1 + MPAD);
    if (y1 - y0 < 1 || x1 - x0 < 2) continue;
    for (int y = y0; y < y1; ++y)
      for (int x = x0; x < x1; ++x) {
        const size_t i = size_t(y) * W + x;
        P[i] = graded(st, cl[i], cr[i - d], L[i], R[i - d]);
      }
    // rectangle filter (de_dense.cpp:434-483): H-fwd, H-bwd, V-fwd, V-bwd
    for (int y = y0; y < y1; ++y) {
      int16_t* r = &P[size_t(y) * W];
      const uint16_t* a = &ax[size_t(y) * W];
      int32_t f = r[x0];
      for (int x = x0 + 1; x < x1; ++x) { f = rf_step(f, r[x], a[x]); r[x] = int16_t(f); }
      f = r[x1 - 1];
      for (int x = x1 - 2; x >= x0; --x) { f = rf_step(f, r[x], a[x + 1]); r[x] = int16_t(f); }
    }
    for (int y = y0 + 1; y < y1; ++y)
      for (int x = x0; x < x1; ++x) {
        const size_t i = size_t(y) * W + x;
        P[i] = int16_t(rf_step(P[i - W], P[i], ay[i]));
      }
    for (int y = y1 - 2; y >= y0; --y)
      for (int x = x0; x < x1; ++x) {
        const size_t i = size_t(y) * W + x;
        P[i] = int16_t(rf_step(P[i + W], P[i], ay[i + W]));
      }
    const int16_t kk = int16_t(k);
    for (int y = std::max(y0, pylo[k]); y < std::min(y1, pyhi[k] + 1); ++y) {
      const int xa = std::max(x0, int(mx0[size_t(y) * D + k]));
      const int xb = std::min(x1, int(mx1[size_t(y) * D + k]) + 1);
      for (int x = xa; x < xb; ++x) {
        const size_t i = size_t(y) * W + x;
        const int16_t v = P[i];
        if (cfg.lrc) {
          const size_t j = i - size_t(d);
          if (v > rs[j]) { rs[j] = v; rd[j] = kk; }
        }
        if (v <= s1[i]) continue;
        if (v > s0[i]) { s1[i] = s0[i]; d1[i] = d0[i]; s0[i] = v; d0[i] = kk; }
        else { s1[i] = v; d1[i] = kk; }
      }
    }
  }
  finish_dense(W, H, cfg, s0, s1, d0, d1, m0, p0, rd, disp, margin, cand6, tie_mask);
  return 0;
}

// Nearest upsample of a half-resolution prior (a disparity d at half scale is
// 2d at full scale), de_dense.cpp:1789-1795 and 1958-1962.
void orc_prior_upsample(const float* half, int W, int H, float* full) {
  const int W2 = W / 2, H2 = H / 2;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x)
      full[size_t(y) * W + x] =
          2.f * half[size_t(std::min(H2 - 1, y / 2)) * W2 + std::min(W2 - 1, x / 2)];
}

// --c2f (de_dense.cpp:1889-1966): a half-resolution pass with D/2 planes, one
// iteration and min_margin 0; holes filled from row neighbours (forward, then
// backward), then column neighbours; the upsampled result is the prior of the
// masked full-resolution sweep. If holes remain, the full sweep runs instead.
// `coarse` (nullable) receives the filled W/2 x H/2 coarse disparity.
int orc_dense_run_c2f(const uint8_t* L, const uint8_t* R, int W, int H, const OrcCfg* cfgp,
                      int band, float* disp, float* margin, float* cand6, uint8_t* tie_mask,
                      float* coarse, int* used_mask) {
  if (!L || !R || !cfgp || !disp || !margin || W < 2 || H < 2) return -1;
  OrcCfg cfg = *cfgp;
  cfg.subpixel = 0;
  const int W2 = W / 2, H2 = H / 2;
  std::vector<uint8_t> L2(size_t(W2) * H2), R2(size_t(W2) * H2);
  for (int y = 0; y < H2; ++y)
    for (int x = 0; x < W2; ++x) {
      const size_t a = size_t(2 * y) * W + 2 * x;
      L2[size_t(y) * W2 + x] = uint8_t((int(L[a]) + L[a + 1] + L[a + W] + L[a + W + 1] + 2) / 4);
      R2[size_t(y) * W2 + x] = uint8_t((int(R[a]) + R[a + 1] + R[a + W] + R[a + W + 1] + 2) / 4);
    }
  OrcCfg cc = cfg;
  cc.dmin = std::max(1, cfg.dmin / 2);
  cc.dmax = (cfg.dmax + 1) / 2;
  cc.min_margin = 0.f;
  cc.iters = 1;
  std::vector<float> dh(size_t(W2) * H2), mh(size_t(W2) * H2);
  if (int rc = orc_dense_run(L2.data(), R2.data(), W2, H2, &cc, dh.data(), mh.data(), nullptr,
                             nullptr))
    return rc;
  for (int y = 0; y < H2; ++y) {
    float* r = &dh[size_t(y) * W2];
    float last = std::nanf("");
    for (int x = 0; x < W2; ++x) {
      if (std::isfinite(r[x])) last = r[x];
      else if (std::isfinite(last)) r[x] = last;
    }
    last = std::nanf("");
    for (int x = W2 - 1; x >= 0; --x) {
      if (std::isfinite(r[x])) last = r[x];
      else if (std::isfinite(last)) r[x] = last;
    }
  }
  for (int x = 0; x < W2; ++x) {
    float last = std::nanf("");
    for (int y = 0; y < H2; ++y) {
      float& v = dh[size_t(y) * W2 + x];
      if (std::isfinite(v)) last = v;
      else if (std::isfinite(last)) v = last;
    }
    last = std::nanf("");
    for (int y = H2 - 1; y >= 0; --y) {
      float& v = dh[size_t(y) * W2 + x];
      if (std::isfinite(v)) last = v;
      else if (std::isfinite(last)) v = last;
    }
  }
  if (coarse) std::copy(dh.begin(), dh.end(), coarse);
  size_t holes = 0;
  for (float v : dh) holes += std::isfinite(v) ? 0 : 1;
  if (used_mask) *used_mask = holes ? 0 : 1;
  if (holes) return orc_dense_run(L, R, W, H, &cfg, disp, margin, cand6, tie_mask);
  std::vector<float> prior(size_t(W) * H);
  orc_prior_upsample(dh.data(), W, H, prior.data());
  return orc_dense_run_prior(L, R, W, H, &cfg, prior.data(), band, disp, margin, cand6,
                             tie_mask);
}

// Aggregated FP32 volume of the full-volume path (--no-blockwise / --dump-vol):
// census-only float score (half-h)/half, float recursive filter, -1e30 outside
// the valid window. Layout [y][x][k] like the dump. core/tools/de_dense.cpp:
// 1305-1318 (score), 510-527 (filter), 1395-1411 (stage + transpose).
int orc_volume_f32(const uint8_t* L, const uint8_t* R, int W, int H,
                   const OrcCfg* cfg, float* vol) {
  const int hw = cfg->census_hw, hh = cfg->census_hh;
  const int nb = (2 * hw + 1) * (2 * hh + 1) - 1;
  if (nb > 64 || (nb & 1)) return -2;
  const int D = cfg->dmax - cfg->dmin + 1;
  const size_t WH = size_t(W) * H;
  std::vector<uint64_t> cl(WH), cr(WH);
  census_plane(L, W, H, hw, hh, cl.data());
  census_plane(R, W, H, hw, hh, cr.data());
  std::vector<uint16_t> ax(WH), ay(WH);
  orc_rf_coeffs(L, W, H, cfg->sigma_s, cfg->sigma_r, ax.data(), ay.data());
  const float half = float(nb / 2);
  const float inv = 1.f / half;
  std::vector<float> P(WH);
  std::fill(vol, vol + WH * size_t(D), kNone);
  for (int k = 0; k < D; ++k) {
    const int d = cfg->dmin + k;
    std::fill(P.begin(), P.end(), 0.f);
    for (int y = hh; y < H - hh; ++y)
      for (int x = hw + d; x < W - hw; ++x) {
        const size_t i = size_t(y) * W + x;
        P[i] = (half - float(__builtin_popcountll(cl[i] ^ cr[i - d]))) * inv;
      }
    rf_filter_plane_f32(P.data(), ax.data(), ay.data(), W, H);
    for (int y = hh; y < H - hh; ++y)
      for (int x = hw + d; x < W - hw; ++x) {
        const size_t i = size_t(y) * W + x;
        vol[i * D + k] = P[i];
      }
  }
  return 0;
}

// Full-volume row solver (--topk 0): scores over each pixel's whole valid
// disparity interval, `iters` rounds of rho (over a left pixel's disparities)
// and beta (over a right pixel's claimants, the diagonal x = xr + dmin + k),
// best belief per pixel, greedy claim in std::sort order, parabola sub-pixel.
// Follows core/tools/de_dense.cpp:566-712; the caller passes hw = 3 there
// (de_dense.cpp:1541), whatever the census window.
static void solve_row_full(int W, int D, const OrcCfg& cfg, const float* vol, float* s,
                           float* beta, float* rho, float* out_disp, float* out_margin,
                           int* tie_flag) {
  const int hw = 3;
  const int dmin = cfg.dmin;
  const float keep = 1.f - cfg.damping;
  std::vector<int> k0(W), k1(W), bestk(W), order(W);
  std::vector<float> best(W);
  std::vector<char> taken(W);
  std::vector<float> gath(size_t(D) + 2), sv(size_t(D) + 2);
  std::vector<int> ok(size_t(D) + 2);
  for (int x = 0; x < W; ++x) {
    int lo = 0, hi = 0;
    if (x >= hw && x < W - hw) {
      lo = std::max(0, x - (W - hw) + 1 - dmin + 1);
      hi = std::min(D, x - hw - dmin + 1);
      if (hi < lo) hi = lo;
    }
    k0[x] = lo;
    k1[x] = hi;
    Top2f t;
    for (int k = lo; k < hi; ++k) {
      const float v = vol[size_t(x) * D + k];
      s[size_t(x) * D + k] = v;
      if (v > -1e29f) t.push(v, k);
    }
    const float alt = (t.b2 < -1e29f) ? cfg.lambda : std::max(t.b2, cfg.lambda);
    out_margin[x] = (t.i1 < 0) ? 0.f : t.b1 - alt;
  }
  std::fill(beta, beta + size_t(W) * D, 0.f);
  std::fill(rho, rho + size_t(W) * D, 0.f);
  for (int it = 0; it < cfg.iters; ++it) {
    for (int x = 0; x < W; ++x) {
      const int lo = k0[x], hi = k1[x];
      if (hi <= lo) continue;
      Top2f t;
      for (int k = lo; k < hi; ++k) t.push(beta[size_t(x) * D + k], k);
      for (int k = lo; k < hi; ++k) {
        const size_t i = size_t(x) * D + k;
        const float tgt = s[i] - std::max(cfg.lambda, t.excl(k));
        rho[i] = keep * tgt + cfg.damping * rho[i];
      }
    }
    for (int xr = 0; xr < W; ++xr) {
      const int kmax = std::min(D, W - xr - dmin);
      if (kmax <= 0) break;
      Top2f t;
      int nv = 0;
      for (int k = 0; k < kmax; ++k) {
        const int x = xr + dmin + k;
        ok[size_t(k)] = (k >= k0[x] && k < k1[x]);
        if (!ok[size_t(k)]) continue;
        const size_t i = size_t(x) * D + k;
        gath[size_t(k)] = rho[i];
        sv[size_t(k)] = s[i];
        t.push(gath[size_t(k)], k);
        ++nv;
      }
      if (nv == 0) continue;
      for (int k = 0; k < kmax; ++k) {
        if (!ok[size_t(k)]) continue;
        const size_t i = size_t(xr + dmin + k) * D + k;
        beta[i] = keep * (sv[size_t(k)] - std::max(cfg.gamma, t.excl(k))) + cfg.damping * beta[i];
      }
    }
  }
  for (int x = 0; x < W; ++x) {
    best[size_t(x)] = -1e30f;
    bestk[size_t(x)] = -1;
    taken[size_t(x)] = 0;
    order[size_t(x)] = x;
    for (int k = k0[x]; k < k1[x]; ++k) {
      const size_t i = size_t(x) * D + k;
      if (s[i] <= cfg.lambda) continue;
      const float bel = beta[i] + rho[i] - s[i];
      if (bel > best[size_t(x)]) { best[size_t(x)] = bel; bestk[size_t(x)] = k; }
    }
  }
  if (tie_flag) {
    std::vector<float> top(W, kNone);
    std::vector<int> cnt(W, 0);
    for (int x = 0; x < W; ++x) {
      tie_flag[x] = 0;
      if (bestk[size_t(x)] < 0) continue;
      const int xr = x - (dmin + bestk[size_t(x)]);
      if (xr < 0) continue;
      if (best[size_t(x)] > top[size_t(xr)]) { top[size_t(xr)] = best[size_t(x)]; cnt[size_t(xr)] = 1; }
      else if (best[size_t(x)] == top[size_t(xr)]) ++cnt[size_t(xr)];
    }
    for (int x = 0; x < W; ++x) {
      if (bestk[size_t(x)] < 0) continue;
      const int xr = x - (dmin + bestk[size_t(x)]);
      if (xr >= 0 && cnt[size_t(xr)] > 1 && best[size_t(x)] == top[size_t(xr)]) tie_flag[x] = 1;
    }
  }
  std::sort(order.begin(), order.end(), [&](int a, int b) { return best[size_t(a)] > best[size_t(b)]; });
  for (int oi = 0; oi < W; ++oi) {
    const int x = order[size_t(oi)];
    out_disp[x] = std::nanf("");
    if (bestk[size_t(x)] < 0) continue;
    if (cfg.min_margin > 0.f && out_margin[x] < cfg.min_margin) continue;
    const int k = bestk[size_t(x)];
    const int d = dmin + k;
    const int xr = x - d;
    if (xr < 0 || taken[size_t(xr)]) continue;
    taken[size_t(xr)] = 1;
    float off = 0.f;
    const float* sp = s + size_t(x) * D;
    if (cfg.subpixel && k > k0[x] && k + 1 < k1[x]) {
      const float cm = sp[k - 1], c0 = sp[k], cp = sp[k + 1];
      const float den = 2.f * c0 - cm - cp;
      if (den > 1e-6f) {
        off = 0.5f * (cp - cm) / den;
        if (off > 0.5f) off = 0.5f;
        if (off < -0.5f) off = -0.5f;
      }
    }
    out_disp[x] = float(d) + off;
  }
}

// --topk 0: FP32 volume + full-volume row solver. tie_mask as orc_dense_run.
int orc_dense_run_full(const uint8_t* L, const uint8_t* R, int W, int H, const OrcCfg* cfgp,
                       float* disp, float* margin, uint8_t* tie_mask) {
  if (!L || !R || !cfgp || !disp || !margin || W <= 0 || H <= 0) return -1;
  const OrcCfg& cfg = *cfgp;
  const int D = cfg.dmax - cfg.dmin + 1;
  if (D < 1) return -3;
  std::vector<float> vol(size_t(W) * H * D);
  if (int rc = orc_volume_f32(L, R, W, H, cfgp, vol.data())) return rc;
  std::vector<float> s(size_t(W) * D), beta(size_t(W) * D), rho(size_t(W) * D);
  std::vector<int> tie(W);
  for (int y = 0; y < H; ++y) {
    solve_row_full(W, D, cfg, &vol[size_t(y) * W * D], s.data(), beta.data(), rho.data(),
                   &disp[size_t(y) * W], &margin[size_t(y) * W], tie_mask ? tie.data() : nullptr);
    if (tie_mask)
      for (int x = 0; x < W; ++x) tie_mask[size_t(y) * W + x] = uint8_t(tie[size_t(x)]);
  }
  return 0;
}

// dense_solve.hpp:158-171.
float orc_confidence(float s1, float s2, float lambda, float lrc, float lrc_tol) {
  if (!(s1 > -1e29f)) return 0.f;
  const float alt = (s2 > -1e29f) ? std::max(s2, lambda) : lambda;
  const float x = std::log(std::max((1.f - alt) / std::max(1.f - s1, 1e-6f), 1e-6f));
  const float logit = 0.692684f + 7.936751f * x;
  const float p = 1.f / (1.f + std::exp(-std::max(-30.f, std::min(30.f, logit))));
  return (lrc > lrc_tol) ? std::min(p, 0.35f) : p;
}

uint8_t orc_confidence_u8(float s1, float s2, float lambda, float lrc, float lrc_tol) {
  return uint8_t(std::lround(orc_confidence(s1, s2, lambda, lrc, lrc_tol) * 255.f));
}

// The DEDD output planes of one frame (core/tools/de_dense_cuda.cu:731-745):
// Q4 disparity lround(16 d) where d > 0 and finite, else -32768; u8 confidence
// of (s1, s2, lambda, lrc if finite else 0), 0 where Q4 is -32768. p255 (nullable)
// receives the float P * 255 the byte rounds, 0 where there is no answer.
void orc_dedd_planes(const float* disp, const float* s1, const float* s2, const float* lrc,
                     int n, float lambda, int16_t* q4, uint8_t* conf, float* p255) {
  for (int i = 0; i < n; ++i) {
    const float d = disp[i];
    q4[i] = (d > 0.f && std::isfinite(d)) ? int16_t(std::lround(d * 16.0f)) : int16_t(-32768);
    const float l = std::isfinite(lrc[i]) ? lrc[i] : 0.f;
    const float p = orc_confidence(s1[i], s2[i], lambda, l, 1.f) * 255.f;
    conf[i] = q4[i] == int16_t(-32768) ? uint8_t(0) : uint8_t(std::lround(p));
    if (p255) p255[i] = q4[i] == int16_t(-32768) ? 0.f : p;
  }
}

}  // extern "C"
