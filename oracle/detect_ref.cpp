// CPU restatement of the reference's FAST keypoint detector.
// TEST INFRASTRUCTURE ONLY: the checker for the HIP detector, never shipped.
//
// Follows core/src/preproc.cpp:
//   is_fast_corner          preproc.cpp:371-399 (FAST-9 on the radius-3 ring)
//   shi_tomasi_at           preproc.cpp:401-427 (3x3 Sobel / 8, box of 2r+1)
//   local_std_at            preproc.cpp:164-181 (double sums)
//   detect_keypoints_fast   preproc.cpp:436-545 (candidates, greedy NMS in
//                           response order, texture floor, sub-pixel parabola,
//                           per-cell cap)
// and core/src/match.cpp:768-810 (refine_disparity: SAD parabola, +-0.5 clamp).
// `canonical` = 0 orders equal responses as the reference's std::sort /
// std::partial_sort do; 1 breaks them by scan order (y, then x), as the HIP
// detector does.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "sparse_ref.h"

namespace {

const int kRingX[16] = {0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3, -3, -3, -2, -1};
const int kRingY[16] = {-3, -3, -2, -1, 0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3};

bool run_of(unsigned mask, int n) {
  const unsigned twice = mask | (mask << 16);
  const unsigned want = (1u << n) - 1u;
  for (int i = 0; i < 16; ++i)
    if (((twice >> i) & want) == want) return true;
  return false;
}

bool fast_corner(const uint8_t* img, int W, int H, int x, int y, int thr) {
  if (x < 3 || y < 3 || x >= W - 3 || y >= H - 3) return false;
  const int c = img[size_t(y) * W + x];
  int bright = 0, dark = 0;
  for (int k = 0; k < 16; k += 4) {
    const int v = img[size_t(y + kRingY[k]) * W + x + kRingX[k]];
    if (v > c + thr) ++bright;
    else if (v < c - thr) ++dark;
  }
  if (bright < 2 && dark < 2) return false;
  unsigned bm = 0, dm = 0;
  for (int k = 0; k < 16; ++k) {
    const int v = img[size_t(y + kRingY[k]) * W + x + kRingX[k]];
    if (v > c + thr) bm |= 1u << k;
    else if (v < c - thr) dm |= 1u << k;
  }
  return run_of(bm, 9) || run_of(dm, 9);
}

float shi_tomasi(const uint8_t* img, int W, int H, int x, int y, int grad_window) {
  const int r = std::max(1, grad_window / 2);
  if (x - r - 1 < 0 || y - r - 1 < 0 || x + r + 1 >= W || y + r + 1 >= H) return 0.f;
  float sxx = 0.f, syy = 0.f, sxy = 0.f;
  for (int dy = -r; dy <= r; ++dy)
    for (int dx = -r; dx <= r; ++dx) {
      const int px = x + dx, py = y + dy;
      const uint8_t* rm = img + size_t(py - 1) * W;
      const uint8_t* r0 = img + size_t(py) * W;
      const uint8_t* rp = img + size_t(py + 1) * W;
      const float gx = (float(rm[px + 1]) - float(rm[px - 1])
                        + 2.f * (float(r0[px + 1]) - float(r0[px - 1]))
                        + float(rp[px + 1]) - float(rp[px - 1])) * 0.125f;
      const float gy = (float(rp[px - 1]) - float(rm[px - 1])
                        + 2.f * (float(rp[px]) - float(rm[px]))
                        + float(rp[px + 1]) - float(rm[px + 1])) * 0.125f;
      sxx += gx * gx;
      syy += gy * gy;
      sxy += gx * gy;
    }
  const float n = 1.f / float((2 * r + 1) * (2 * r + 1));
  const float a = sxx * n, b = syy * n, c = sxy * n;
  const float half_tr = 0.5f * (a + b);
  const float disc = std::sqrt(std::max(0.f, 0.25f * (a - b) * (a - b) + c * c));
  return half_tr - disc;
}

float local_std7(const uint8_t* img, int W, int H, int x, int y) {
  const int r = 3;
  const int x0 = std::max(0, x - r), x1 = std::min(W - 1, x + r);
  const int y0 = std::max(0, y - r), y1 = std::min(H - 1, y + r);
  double s1 = 0.0, s2 = 0.0;
  for (int yy = y0; yy <= y1; ++yy)
    for (int xx = x0; xx <= x1; ++xx) {
      const double v = img[size_t(yy) * W + xx];
      s1 += v;
      s2 += v * v;
    }
  const double n = double((y1 - y0 + 1) * (x1 - x0 + 1));
  const double mean = s1 / n;
  const double var = s2 / n - mean * mean;
  return var > 0.0 ? float(std::sqrt(var)) : 0.f;
}

}  // namespace

extern "C" {

void orc_detector_default_cfg(OrcDetectorCfg* c) {
  // DetectorConfig defaults (core/include/doubleeye/preproc.hpp:132-145)
  c->grad_window = 3;
  c->cell = 32;
  c->per_cell = 3;
  c->min_response = 1.0f;
  c->min_local_std = 2.0f;
  c->border = 8;
  c->subpixel = 1;
  c->fast_threshold = 8;
  c->nms_radius = 3;
}

int orc_is_fast_corner(const uint8_t* img, int W, int H, int x, int y, int thr) {
  return fast_corner(img, W, H, x, y, thr) ? 1 : 0;
}

float orc_shi_tomasi_at(const uint8_t* img, int W, int H, int x, int y, int grad_window) {
  return shi_tomasi(img, W, H, x, y, grad_window);
}

float orc_local_std_at(const uint8_t* img, int W, int H, int x, int y) {
  return local_std7(img, W, H, x, y);
}

int orc_detect_keypoints_fast(const uint8_t* img, int W, int H, const OrcDetectorCfg* cfg,
                              int canonical, OrcKeypoint* out, int cap, int* n_out, int* counts) {
  *n_out = 0;
  if (counts) counts[0] = counts[1] = 0;
  if (W <= 0 || H <= 0) return 0;
  const OrcDetectorCfg& c = *cfg;
  const int border = std::max(c.border, 5);
  struct Cand {
    OrcKeypoint kp;
    int scan;
  };
  std::vector<Cand> cand;
  for (int y = border; y < H - border; ++y)
    for (int x = border; x < W - border; ++x) {
      if (!fast_corner(img, W, H, x, y, c.fast_threshold)) continue;
      Cand k;
      k.kp.x = float(x);
      k.kp.y = float(y);
      k.kp.response = shi_tomasi(img, W, H, x, y, c.grad_window);
      k.kp.local_std = 0.f;
      k.scan = y * W + x;
      if (k.kp.response < c.min_response) continue;
      cand.push_back(k);
    }
  if (counts) counts[0] = int(cand.size());
  if (cand.empty()) return 0;
  auto stronger = [](const Cand& a, const Cand& b) { return a.kp.response > b.kp.response; };
  if (canonical) std::stable_sort(cand.begin(), cand.end(), stronger);
  else std::sort(cand.begin(), cand.end(), stronger);
  const int rad = std::max(1, c.nms_radius);
  std::vector<uint8_t> taken(size_t(W) * H, 0);
  std::vector<Cand> kept;
  for (const Cand& k : cand) {
    const int xi = int(k.kp.x), yi = int(k.kp.y);
    bool blocked = false;
    for (int dy = -rad; dy <= rad && !blocked; ++dy) {
      const int py = yi + dy;
      if (py < 0 || py >= H) continue;
      for (int dx = -rad; dx <= rad; ++dx) {
        const int px = xi + dx;
        if (px < 0 || px >= W) continue;
        if (taken[size_t(py) * W + px]) { blocked = true; break; }
      }
    }
    if (blocked) continue;
    taken[size_t(yi) * W + xi] = 1;
    kept.push_back(k);
  }
  if (counts) counts[1] = int(kept.size());
  const int cell = std::max(4, c.cell);
  const int cols = (W + cell - 1) / cell, rows = (H + cell - 1) / cell;
  std::vector<std::vector<Cand>> buckets(size_t(cols) * rows);
  for (Cand k : kept) {
    const int xi = int(k.kp.x), yi = int(k.kp.y);
    k.kp.local_std = local_std7(img, W, H, xi, yi);
    if (k.kp.local_std < c.min_local_std) continue;
    if (c.subpixel) {
      const float r0 = k.kp.response;
      const float l = shi_tomasi(img, W, H, xi - 1, yi, c.grad_window);
      const float r = shi_tomasi(img, W, H, xi + 1, yi, c.grad_window);
      const float dxx = l - 2.f * r0 + r;
      if (std::fabs(dxx) > 1e-12f) {
        const float d = 0.5f * (l - r) / dxx;
        if (std::fabs(d) <= 0.5f) k.kp.x += d;
      }
      const float u = shi_tomasi(img, W, H, xi, yi - 1, c.grad_window);
      const float dn = shi_tomasi(img, W, H, xi, yi + 1, c.grad_window);
      const float dyy = u - 2.f * r0 + dn;
      if (std::fabs(dyy) > 1e-12f) {
        const float d = 0.5f * (u - dn) / dyy;
        if (std::fabs(d) <= 0.5f) k.kp.y += d;
      }
    }
    buckets[size_t(yi / cell) * cols + (xi / cell)].push_back(k);
  }
  const size_t keep = size_t(std::max(1, c.per_cell));
  int n = 0;
  for (auto& b : buckets) {
    if (b.size() > keep) {
      if (canonical) std::stable_sort(b.begin(), b.end(), stronger);
      else std::partial_sort(b.begin(), b.begin() + long(keep), b.end(), stronger);
      b.resize(keep);
    }
    for (const Cand& k : b) {
      if (n >= cap) return -1;
      out[n++] = k.kp;
    }
  }
  *n_out = n;
  return 0;
}

void orc_refine_disparity(const uint8_t* L, const uint8_t* R, int W, int H,
                          const OrcKeypoint* kl, int nl, const OrcKeypoint* kr, int nr,
                          OrcMatch* m, int n, int half) {
  (void)kr;   // right keypoints only bound the match's right index
  if (W <= 0 || H <= 0) return;
  auto cost = [&](int xl, int y, int d) -> long {
    const int xr = xl - d;
    if (xl - half < 0 || xl + half >= W) return -1;
    if (xr - half < 0 || xr + half >= W) return -1;
    if (y - half < 0 || y + half >= H) return -1;
    long acc = 0;
    for (int v = -half; v <= half; ++v)
      for (int u = -half; u <= half; ++u)
        acc += std::abs(int(L[size_t(y + v) * W + xl + u]) - int(R[size_t(y + v) * W + xr + u]));
    return acc;
  };
  for (int k = 0; k < n; ++k) {
    OrcMatch& q = m[k];
    if (q.left < 0 || q.right < 0 || q.left >= nl || q.right >= nr) continue;
    const int xl = int(kl[q.left].x + 0.5f);
    const int y = int(kl[q.left].y + 0.5f);
    const int d0 = int(std::lround(q.disparity));
    const long c0 = cost(xl, y, d0), cm = cost(xl, y, d0 - 1), cp = cost(xl, y, d0 + 1);
    if (c0 < 0 || cm < 0 || cp < 0) continue;
    const double denom = double(cm) - 2.0 * double(c0) + double(cp);
    if (denom <= 0.0) continue;
    double off = 0.5 * (double(cm) - double(cp)) / denom;
    off = std::min(0.5, std::max(-0.5, off));
    q.disparity = float(double(d0) + off);
  }
}

}  // extern "C"
