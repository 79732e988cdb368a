// Oracle (test infrastructure, not product code): multi-scale output resampling in FP64
// (RenderFrameToScreen, libs/vis_utils/renderoutputframe.cpp:243-535 and the shaders
// vis_utils/shader/renderoutputframe/*.comp; kernels and digital filters of
// vis_utils/filters/*.hpp, after Nehab and Hoppe 2014).  Frames are RGBA, row-major,
// 4 doubles per pixel; samples outside the frame read 0 (texelFetch out of range).
#pragma once
#include <cmath>
#include <vector>

namespace orc {

enum { K_BOX = 0, K_HAT = 1, K_CATMULL_ROM = 2, K_MITCHELL = 3, K_BSPLINE3 = 4, K_OMOMS3 = 5 };
enum { S_MULTISAMPLE = 0, S_DOWNSCALE = 1, S_UPSCALE = 2 };

// Polynomial pieces k0 (|x| in (1, 2], u = 2 - |x|) and k1 (|x| in [0, 1], u = 1 - |x|):
// catmullrom_filter.comp, mitchellnetravali_filter.comp, cardinalbspline_filter.comp,
// cardinalomoms_filter.comp.
inline double kernel_piece(int k, double u, bool outer) {
  switch (k) {
    case K_CATMULL_ROM: return outer ? ((.5 * u - .5) * u) * u : ((-1.5 * u + 2.0) * u + .5) * u;
    case K_MITCHELL:
      return outer ? (((7 / 18.0) * u - 1 / 3.0) * u) * u : (((-7 / 6.0) * u + 1.5) * u + 0.5) * u + 1 / 18.0;
    case K_BSPLINE3: return outer ? (u * u) * u : ((-3.0 * u + 3.0) * u + 3.0) * u + 1.0;
    default: return outer ? ((.875 * u) * u + .125) * u : ((-2.625 * u + 2.625) * u + 2.25) * u + 1.0;
  }
}
inline double kernel_support(int k) { return k == K_BOX ? 1.0 : (k == K_HAT ? 2.0 : 4.0); }
inline double kernel_weight(int k, double x) {
  if (k == K_BOX) return (x <= -0.5 || x > 0.5) ? 0.0 : 1.0;  // box_filter.comp
  x = std::fabs(x);
  if (k == K_HAT) return x > 1.0 ? 0.0 : 1.0 - x;           // hat_filter.comp
  return x > 2.0 ? 0.0 : (x > 1.0 ? kernel_piece(k, 2.0 - x, true) : kernel_piece(k, 1.0 - x, false));
}

struct Frame {
  int w = 0, h = 0;
  std::vector<double> p;  // 4 per pixel
  double at(int c, int r, int ch) const {
    if (c < 0 || r < 0 || c >= w || r >= h) return 0.0;
    return p[(size_t(r) * w + c) * 4 + ch];
  }
};

// multisample_filter.comp: GL_LINEAR / CLAMP_TO_EDGE at the target pixel centres.
inline void resample_multisample(const Frame& s, Frame& d) {
  for (int y = 0; y < d.h; ++y)
    for (int x = 0; x < d.w; ++x) {
      const double u = (x + 0.5) / d.w * s.w - 0.5, v = (y + 0.5) / d.h * s.h - 0.5;
      const double fu = std::floor(u), fv = std::floor(v);
      const double ax = u - fu, ay = v - fv;
      auto cl = [](long long i, int n) { return int(std::min<long long>(std::max<long long>(i, 0), n - 1)); };
      const int x0 = cl((long long)fu, s.w), x1 = cl((long long)fu + 1, s.w);
      const int y0 = cl((long long)fv, s.h), y1 = cl((long long)fv + 1, s.h);
      for (int c = 0; c < 4; ++c) {
        const double a = s.at(x0, y0, c) * (1 - ax) + s.at(x1, y0, c) * ax;
        const double b = s.at(x0, y1, c) * (1 - ax) + s.at(x1, y1, c) * ax;
        d.p[(size_t(y) * d.w + x) * 4 + c] = a * (1 - ay) + b * ay;
      }
    }
}

// downscaling_filter.comp:17-57
inline void resample_down(const Frame& s, Frame& d, int k) {
  const double kr = 0.5 * kernel_support(k), s_r = double(d.h) / s.h, s_c = double(d.w) / s.w;
  for (int jr = 0; jr < d.h; ++jr)
    for (int jc = 0; jc < d.w; ++jc) {
      const double x_r = (jr + 0.5) / d.h, x_c = (jc + 0.5) / d.w;
      const int il_r = int(std::ceil((x_r - kr / d.h) * s.h - 0.5)), ir_r = int(std::floor((x_r + kr / d.h) * s.h - 0.5));
      const int il_c = int(std::ceil((x_c - kr / d.w) * s.w - 0.5)), ir_c = int(std::floor((x_c + kr / d.w) * s.w - 0.5));
      double f[4] = {0, 0, 0, 0};
      for (int i_r = il_r; i_r <= ir_r; ++i_r)
        for (int i_c = il_c; i_c <= ir_c; ++i_c) {
          const double w = kernel_weight(k, (x_r - (i_r + 0.5) / s.h) * d.h) * kernel_weight(k, (x_c - (i_c + 0.5) / s.w) * d.w);
          for (int c = 0; c < 4; ++c) f[c] += w * s.at(i_c, i_r, c);
        }
      for (int c = 0; c < 4; ++c) d.p[(size_t(jr) * d.w + jc) * 4 + c] = f[c] * (s_r * s_c);
    }
}

// upscaling_filter.comp:17-55
inline void resample_up(const Frame& s, Frame& d, int k) {
  const double kr = 0.5 * kernel_support(k);
  for (int jr = 0; jr < d.h; ++jr)
    for (int jc = 0; jc < d.w; ++jc) {
      const double xi_r = (jr + 0.5) / d.h * s.h - 0.5, xi_c = (jc + 0.5) / d.w * s.w - 0.5;
      const int il_r = int(std::ceil(xi_r - kr)), ir_r = int(std::floor(xi_r + kr));
      const int il_c = int(std::ceil(xi_c - kr)), ir_c = int(std::floor(xi_c + kr));
      double f[4] = {0, 0, 0, 0};
      for (int i_r = il_r; i_r <= ir_r; ++i_r)
        for (int i_c = il_c; i_c <= ir_c; ++i_c) {
          const double w = kernel_weight(k, xi_r - i_r) * kernel_weight(k, xi_c - i_c);
          for (int c = 0; c < 4; ++c) f[c] += w * s.at(i_c, i_r, c);
        }
      for (int c = 0; c < 4; ++c) d.p[(size_t(jr) * d.w + jc) * 4 + c] = f[c];
    }
}

// linear_solve (filters/utils.hpp:17-44) on one channel sequence, with the LU factors of
// the sampled kernel (1, a, 1): L_0 = 1/(a+1), L_i = 1/(a - L_{i-1}), m factors.
inline void linear_solve(std::vector<double>& f, int k) {
  const double a = k == K_BSPLINE3 ? 4.0 : 3.25;
  const int m = k == K_BSPLINE3 ? 8 : 9;
  std::vector<double> L(m);
  double l = 1.0 / (a + 1.0);
  for (int i = 0; i < m; ++i) { L[i] = double(float(l)); l = 1.0 / (a - l); }
  const size_t n = f.size();
  while (f.size() <= size_t(m)) {
    std::vector<double> r(f.rbegin(), f.rend());
    f.insert(f.end(), r.begin(), r.end());
  }
  const int nn = int(f.size());
  const double Li = L[m - 1], v_inv = Li / (1.0 + Li);
  for (int i = 1; i < m; ++i) f[i] -= L[i - 1] * f[i - 1];
  for (int i = m; i < nn; ++i) f[i] -= Li * f[i - 1];
  f[nn - 1] *= v_inv;
  for (int i = nn - 2; i >= m - 1; --i) f[i] = Li * (f[i] - f[i + 1]);
  for (int i = m - 2; i >= 0; --i) f[i] = L[i] * (f[i] - f[i + 1]);
  f.resize(n);
}

// digital_filter2D: every row, then every column, per channel.
inline void digital_filter(Frame& g, int k) {
  std::vector<double> seq;
  for (int c = 0; c < 4; ++c) {
    for (int r = 0; r < g.h; ++r) {
      seq.resize(g.w);
      for (int x = 0; x < g.w; ++x) seq[x] = g.p[(size_t(r) * g.w + x) * 4 + c];
      linear_solve(seq, k);
      for (int x = 0; x < g.w; ++x) g.p[(size_t(r) * g.w + x) * 4 + c] = seq[x];
    }
    for (int x = 0; x < g.w; ++x) {
      seq.resize(g.h);
      for (int r = 0; r < g.h; ++r) seq[r] = g.p[(size_t(r) * g.w + x) * 4 + c];
      linear_solve(seq, k);
      for (int r = 0; r < g.h; ++r) g.p[(size_t(r) * g.w + x) * 4 + c] = seq[r];
    }
  }
}

inline void resample(const Frame& src, Frame& dst, int mode, int k) {
  const bool cardinal = k == K_BSPLINE3 || k == K_OMOMS3;
  dst.p.assign(size_t(dst.w) * dst.h * 4, 0.0);
  if (mode == S_MULTISAMPLE) return resample_multisample(src, dst);
  if (mode == S_DOWNSCALE) {
    resample_down(src, dst, k);
    if (cardinal) digital_filter(dst, k);
    return;
  }
  if (!cardinal) return resample_up(src, dst, k);
  Frame pre = src;
  digital_filter(pre, k);
  resample_up(pre, dst, k);
}

}  // namespace orc
