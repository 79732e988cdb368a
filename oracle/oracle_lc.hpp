// Oracle (test infrastructure, not product code): the object-space light cache of the
// EBS and DOS renderers and the march that reads it.
//
// PreIlluminationStructuredVolume (cppvolrend/utils/preillumination.cpp:7-80) holds a
// 2-channel 3D texture (occlusion, shadow), 32^3 by default (ebsrenderer.cpp:46-47,
// dosrcrenderer.cpp:64), filled every Update() by a compute pass over its cells
// (rc1pextbsd/lightcachecomputation.comp:440-468, rc1pdosct/lightcachecomputation.comp:
// 523-546): cell i evaluates the lighting at texture-space position
// (i + 1/2) * scale * (dims / res), channel defaults 1 when the term is off.  The EBS
// cache uses the renderer's own AO and shadow functions; the DOS cache turns the
// occlusion cone towards the eye with a frame built from the camera's up vector
// (:279-291) instead of the camera ray's frame.  The frame is then marched by
// _common_shaders/obj_ray_marching.comp:223-333, which reads the cache with GL_LINEAR
// and CLAMP_TO_EDGE at tex_pos / VolumeScaledSizes and shades
// rgb <- (rgb*Iocc*ka + rgb*Isdw*kd) / (ka + kd) (non-Phong branch, :180-220).
#pragma once
#include "oracle.hpp"
#include "oracle_dos.hpp"

namespace orc {

struct LightCacheO {
  int rx = 0, ry = 0, rz = 0;
  std::vector<double> occ, sdw;  // x fastest
};

// Cell centre in texture space: (i + 1/2) * (scale * (dims / res)), per axis.
inline V3<double> lc_cell_pos(const Volume& v, int rx, int ry, int rz, int x, int y, int z) {
  return V3<double>((x + 0.5) * (v.sx * (double(v.nx) / rx)), (y + 0.5) * (v.sy * (double(v.ny) / ry)),
                    (z + 0.5) * (v.sz * (double(v.nz) / rz)));
}

template <class Cell>
inline void lc_fill(const Volume& v, int rx, int ry, int rz, Cell&& cell, LightCacheO& lc) {
  lc.rx = rx; lc.ry = ry; lc.rz = rz;
  const size_t n = size_t(rx) * ry * rz;
  lc.occ.assign(n, 1.0);
  lc.sdw.assign(n, 1.0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int z = 0; z < rz; ++z)
    for (int y = 0; y < ry; ++y)
      for (int x = 0; x < rx; ++x) {
        const size_t i = size_t(x) + size_t(rx) * (size_t(y) + size_t(ry) * z);
        cell(lc_cell_pos(v, rx, ry, rz, x, y, z), &lc.occ[i], &lc.sdw[i]);
      }
}

// rc1pextbsd/lightcachecomputation.comp:440-468
inline void lc_ebs(const Volume& v, const double* sat, const EbsParams& prm, int rx, int ry, int rz, LightCacheO& lc) {
  Ebs<double, double> e(sat, v, prm);
  lc_fill(v, rx, ry, rz, [&](const V3<double>& tp, double* o, double* s) {
    if (prm.apply_occlusion) *o = e.ambient_occlusion(tp);
    if (prm.apply_shadow) *s = e.shadows(tp, nullptr);
  }, lc);
}

// rc1pdosct/lightcachecomputation.comp:279-291 (occlusion frame) and :523-546
inline void lc_dos(const Volume& v, const Pyramid& P, const DosParamsO& dp, double sigma0, const CamIn& cam, int W,
                   int H, int rx, int ry, int rz, LightCacheO& lc, bool* ok) {
  DosParamsO q = dp;
  for (int i = 0; i < 3; ++i) q.eye[i] = cam.eye[i];
  DosTracer<double> tr(P, q, sigma0);
  *ok = tr.ok;
  if (!tr.ok) return;
  const CamBasis<double> cb = cam_basis<double>(cam, W, H);
  const V3<double> G(v.nx * v.sx, v.ny * v.sy, v.nz * v.sz), eye = cb.eye, cam_up = cb.u;
  lc_fill(v, rx, ry, rz, [&](const V3<double>& tp, double* o, double* s) {
    const V3<double> wp = tp - G * 0.5;
    if (q.apply_occ) {
      const V3<double> k = normalize(eye - wp);
      const V3<double> vr = normalize(cross(-k, cam_up));
      const V3<double> vu = normalize(cross(k, vr));
      *o = dos_cone<double>(P, tr.to, q.occ, tp, k, vu, vr);
    }
    if (q.apply_sdw) *s = tr.shadow(tp, nullptr);
  }, lc);
}

// obj_ray_marching.comp:223-333 with the cache read by GL_LINEAR at tex_pos / G.
template <class R>
inline void render_lc(const Volume& vol, const TF& tf, const LightCacheO& lc, int apply_occ, int apply_sdw, double ka_,
                      double kd_, const CamIn& cam, int W, int H, R step, R et, R* out, double* margin, int stride = 1,
                      const PhongO* ph = nullptr) {
  PhongO Ph = ph ? *ph : PhongO{};
  for (int i = 0; i < 3; ++i) Ph.eye[i] = cam.eye[i];
  CamBasis<R> cb = cam_basis<R>(cam, W, H);
  V3<R> G(R(vol.nx * vol.sx), R(vol.ny * vol.sy), R(vol.nz * vol.sz));
  V3<R> invG(R(1) / G.x, R(1) / G.y, R(1) / G.z);
  const R ka = apply_occ ? R(ka_) : R(0), kd = apply_sdw ? R(kd_) : R(0);
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
        auto sh = [&](const V3<R>& p, R* src) {
          const V3<R> uvw = p * invG;
          const R Io = tex3d_grid<R, double>(lc.occ.data(), lc.rx, lc.ry, lc.rz, uvw);
          const R Is = tex3d_grid<R, double>(lc.sdw.data(), lc.rx, lc.ry, lc.rz, uvw);
          if (ph) return shade_lit_phong(Ph, G, p, Io, Is, ka, kd, apply_sdw ? R(Ph.ks) : R(0), src);  // :186-211
          for (int c = 0; c < 3; ++c) src[c] = (R(1) / (ka + kd)) * (src[c] * Io * ka + src[c] * Is * kd);
        };
        march(tp, d, D, step, et, G, vol, tf.rgbt.data(), tf.size(), sh, o, &mg);
      }
      if (margin) { margin[(size_t(py) * W + px) * 2] = mg.a; margin[(size_t(py) * W + px) * 2 + 1] = mg.g; }
    }
}

}  // namespace orc
