// Oracle (test infrastructure, not product code): the volume gradients and the
// gradient-based Blinn-Phong shading of the plain, EBS, DOS and light-cache marchers.
//
//  - Sobel-Feldman: libs/volvis_utils/utils.cpp:287-350 (the compute-shader form
//    _gradient_shading/sobelfeldman_generator.comp:22-66 is the same sum).
//  - Finite differences: GenerateGradientTexture (utils.cpp:146-240) with its defaults
//    (utils.h:50-56: step 1, no filter, normalised; the NaN of a zero difference -> 0).
//  - The plain march with ShadeBlinnPhong (rc1pass/ray_marching_1p.comp:48-81, 156-250).
// The shading terms themselves (PhongO, phong_terms, shade_*) live in oracle.hpp, next
// to the marchers that call them.
#pragma once
#include "oracle.hpp"

namespace orc {

inline void gradient_sobel(const Volume& v, GradientO& g) {
  g.nx = v.nx; g.ny = v.ny; g.nz = v.nz;
  const size_t n = size_t(v.nx) * v.ny * v.nz;
  g.gx.assign(n, 0); g.gy.assign(n, 0); g.gz.assign(n, 0);
#pragma omp parallel for schedule(static)
  for (int z = 0; z < v.nz; ++z)
    for (int y = 0; y < v.ny; ++y)
      for (int x = 0; x < v.nx; ++x) {
        double sx = 0, sy = 0, sz = 0;
        for (int v1 = -1; v1 <= 1; ++v1)
          for (int v2 = -1; v2 <= 1; ++v2) {
            const double w = 4.0 / std::pow(2.0, std::abs(v1) + std::abs(v2));
            sz += v.norm<double>(x + v1, y + v2, z - 1) * w + v.norm<double>(x + v1, y + v2, z + 1) * -w;
            sy += v.norm<double>(x + v1, y - 1, z + v2) * w + v.norm<double>(x + v1, y + 1, z + v2) * -w;
            sx += v.norm<double>(x - 1, y + v2, z + v1) * w + v.norm<double>(x + 1, y + v2, z + v1) * -w;
          }
        const size_t i = size_t(x) + size_t(v.nx) * (size_t(y) + size_t(v.ny) * z);
        g.gx[i] = sx; g.gy[i] = sy; g.gz[i] = sz;
      }
}

inline void gradient_fd(const Volume& v, GradientO& g) {
  g.nx = v.nx; g.ny = v.ny; g.nz = v.nz;
  const size_t n = size_t(v.nx) * v.ny * v.nz;
  g.gx.assign(n, 0); g.gy.assign(n, 0); g.gz.assign(n, 0);
#pragma omp parallel for schedule(static)
  for (int z = 0; z < v.nz; ++z)
    for (int y = 0; y < v.ny; ++y)
      for (int x = 0; x < v.nx; ++x) {
        V3<double> d(v.norm<double>(x + 1, y, z) - v.norm<double>(x - 1, y, z),
                     v.norm<double>(x, y + 1, z) - v.norm<double>(x, y - 1, z),
                     v.norm<double>(x, y, z + 1) - v.norm<double>(x, y, z - 1));
        const double l = std::sqrt(dot(d, d));
        if (!(l > 0)) d = V3<double>(0, 0, 0);  // normalize(0) is NaN; the reference stores 0
        else d = d * (1.0 / l);
        const size_t i = size_t(x) + size_t(v.nx) * (size_t(y) + size_t(v.ny) * z);
        g.gx[i] = d.x; g.gy[i] = d.y; g.gz[i] = d.z;
      }
}

// plain march with ShadeBlinnPhong (ray_marching_1p.comp:156-250)
template <class R>
inline void render_rc1p_phong(const Volume& vol, const TF& tf, const PhongO& ph, const CamIn& cam, int W, int H, R step,
                              R et, R* out, double* margin, int stride = 1) {
  CamBasis<R> cb = cam_basis<R>(cam, W, H);
  V3<R> G(R(vol.nx * vol.sx), R(vol.ny * vol.sy), R(vol.nz * vol.sz));
  PhongO P = ph;
  for (int i = 0; i < 3; ++i) P.eye[i] = cam.eye[i];
#pragma omp parallel for schedule(dynamic, 1)
  for (int py = 0; py < H; ++py)
    for (int px = 0; px < W; ++px) {
      R* o = out + (size_t(py) * W + px) * 4;
      o[0] = o[1] = o[2] = o[3] = 0;
      if (px % stride || py % stride) continue;
      Margin mg;
      V3<R> d = ray_dir(cb, px, py, W, H);
      R tn, tf_;
      if (ray_aabb(cb.eye, d, G, tn, tf_)) {
        R D = std::fabs(tf_ - tn);
        V3<R> tp = cb.eye + d * tn + G * R(0.5);
        auto sh = [&](const V3<R>& p, R* src) { shade_blinn_phong(P, G, p, src); };
        march(tp, d, D, step, et, G, vol, tf.rgbt.data(), tf.size(), sh, o, &mg);
      }
      if (margin) { margin[(size_t(py) * W + px) * 2] = mg.a; margin[(size_t(py) * W + px) * 2 + 1] = mg.g; }
    }
}

}  // namespace orc
