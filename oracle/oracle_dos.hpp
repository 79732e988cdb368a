// CPU ORACLE — directional occlusion shading with cone section tables and the
// ray-sampled cone ground truth (test infrastructure only).
//
// Section tables: rc1pdosct/conegaussiansampler.cpp:212-414 (FP64, as the reference
// computes them).  Integration: rc1pdosct/ray_bbox_marching.comp:92-112 (lookup),
// :124-333 (occlusion), :345-562 (shadow), :607-656 (shading), :658-734 (march).
// Ground truth: rc1pcrtgt/gt_ray_marching.comp:109-477 with the ray tables of
// crtgtrenderer.cpp:131-185.
//
// The reference stores the section rows in an RGBA16F texture and the ray vectors in
// an RGB16F texture; here (as for every table in this oracle) the values stay at the
// precision the reference's C++ computes them in.
#pragma once
#include "oracle_cone.hpp"
#include <random>

namespace orc {

// ---------------------------------------------------------------- section tables
// ConeGaussianSampler state after its setters' clamps (conegaussiansampler.cpp:57-172).
struct DosConeSpec {
  double half_deg = 20;
  int max_packing = 1;  // 0: _1, 1: _3, 2: _7
  double ui_weight = 0.35;
  double initial_step = 3.0;
  double covered = 100;
  double d_sigma = 1.25, r_sigma = 2.0;
};

struct DosSection {
  int ng;
  double pos, radius, sigma;
  double interval, level, d_integral, amplitude;
};

struct DosTableO {
  std::vector<DosSection> s;
  int n[3] = {0, 0, 0};
  double axes[10][3];  // ray3_axis[0..2], ray7_axis[0..6]
  double adj7 = 0;
};

// conegaussiansampler.h:18-19
static const double kConeDiv3 = 1.0 + 2.0 / std::sqrt(3.0);
static const double kConeDiv7 = 3.01;

// math_utils/utils.cpp:149-165 (RodriguesRotation), in double.
inline V3<double> rodrigues(const V3<double>& v, double t, const V3<double>& k) {
  V3<double> r = v * std::cos(t) + cross(k, v) * std::sin(t) + k * dot(k, v) * (1.0 - std::cos(t));
  return normalize(r);
}

// The three AddGaussianSampleStep* of conegaussiansampler.cpp:418-497, kept as the
// reference chains them: a section tries its current packing first and moves to the
// next packing (permanently) when the cone radius exceeds r_sigma * sigma.
struct DosBuilderO {
  const DosConeSpec& c;
  std::vector<DosSection>& out;
  int ng = 1;
  bool step1(double pos, double sg) {
    if (ng > 1) return step3(pos, sg);
    ng = 1;
    double R = pos * std::tan(c.half_deg * kPi / 180.0);
    if (R > c.r_sigma * sg) return c.max_packing > 0 ? step3(pos, sg) : false;
    out.push_back(DosSection{1, pos, R, sg, 0, 0, 0, 0});
    return true;
  }
  bool step3(double pos, double sg) {
    if (ng > 3) return step7(pos, sg);
    ng = 3;
    double R = pos * std::tan((c.half_deg / kConeDiv3) * kPi / 180.0);
    if (R > c.r_sigma * sg) return c.max_packing > 1 ? step7(pos, sg) : false;
    out.push_back(DosSection{3, pos, R, sg, 0, 0, 0, 0});
    return true;
  }
  bool step7(double pos, double sg) {
    ng = 7;
    double R = pos * std::tan((c.half_deg / kConeDiv7) * kPi / 180.0);
    if (R > c.r_sigma * sg) return false;
    out.push_back(DosSection{7, pos, R, sg, 0, 0, 0, 0});
    return true;
  }
};

// IntegrateGaussian (conegaussiansampler.cpp:393-414): midpoint rule over [-R, R].
inline double dos_integrate_gaussian(double sg, double R) {
  double t = (2.0 * R) / 0.05;
  int nt = int(std::ceil(t));
  double seg = (2.0 * R) / double(nt);
  double s0 = -R + seg * 0.5, S = 0.0;
  for (int i = 0; i < nt; ++i) {
    double x = s0 + seg * double(i);
    S += (1.0 / (std::sqrt(2.0 * kPi) * sg)) * std::exp(-(x * x) / (2.0 * sg * sg)) * seg;
  }
  return S;
}

// ComputeConeIntegrationSteps + ComputeAdditionalInfo (conegaussiansampler.cpp:212-386).
inline bool dos_build(const DosConeSpec& c, double sigma0, DosTableO& T) {
  T.s.clear();
  std::vector<double> ipos, idist;
  DosBuilderO b{c, T.s};
  double pos = c.initial_step, sg = sigma0;
  while (!b.step1(pos, sg)) sg *= 2.0;
  while (pos < c.covered) {
    double si = c.d_sigma * sg;
    while (!b.step1(pos + si + c.d_sigma * sg, sg)) {
      sg *= 2.0;
      if (!(sg < 1e30)) return false;  // the schedule diverges (sigma outgrows every bound)
    }
    si += c.d_sigma * sg;
    ipos.push_back(pos);
    idist.push_back(si);
    pos += si;
  }
  if (idist.empty() || T.s.size() != idist.size() + 1) return false;
  idist.push_back(0.0);
  T.n[0] = T.n[1] = T.n[2] = 0;
  for (size_t i = 0; i < T.s.size(); ++i) {
    DosSection& s = T.s[i];
    T.n[s.ng == 1 ? 0 : (s.ng == 3 ? 1 : 2)]++;
    s.interval = idist[i];
    s.d_integral = i == 0 ? s.sigma * std::sqrt(2.0 * kPi) * 0.5 : idist[i - 1] * 0.5;
    double pr = dos_integrate_gaussian(s.sigma, s.radius);
    double Ac = kPi * s.radius * s.radius;
    double Ig = s.sigma * std::sqrt(2.0 * kPi);
    s.amplitude = ((pr * pr) * (Ig * Ig)) / Ac;
    s.level = std::log2(s.sigma / sigma0);
  }
  // axes (:219-252)
  V3<double> z(0, 0, 1), y(0, 1, 0);
  double t3 = (c.half_deg - c.half_deg / kConeDiv3) * kPi / 180.0;
  V3<double> a3[3];
  a3[0] = rodrigues(rodrigues(z, t3, y), 30.0 * kPi / 180.0, z);
  a3[1] = rodrigues(a3[0], 120.0 * kPi / 180.0, z);
  a3[2] = rodrigues(a3[1], 120.0 * kPi / 180.0, z);
  double t7 = (c.half_deg - c.half_deg / kConeDiv7) * kPi / 180.0;
  V3<double> a7[7];
  a7[0] = z;
  a7[1] = rodrigues(z, t7, y);
  for (int j = 2; j < 7; ++j) a7[j] = rodrigues(a7[j - 1], 60.0 * kPi / 180.0, z);
  for (int j = 0; j < 3; ++j) { T.axes[j][0] = a3[j].x; T.axes[j][1] = a3[j].y; T.axes[j][2] = a3[j].z; }
  for (int j = 0; j < 7; ++j) { T.axes[3 + j][0] = a7[j].x; T.axes[3 + j][1] = a7[j].y; T.axes[3 + j][2] = a7[j].z; }
  T.adj7 = dot(z, a7[1]);
  return true;
}

// ---------------------------------------------------------------- integration
// Cone1/3/7RayOcclusion (ray_bbox_marching.comp:124-322); the shadow versions (:345-531)
// are the same with the shadow table.  Sub-cone i's axis is k*a.z + u*a.y + v*a.x.
template <class R>
inline R dos_cone(const Pyramid& P, const DosTableO& T, const DosConeSpec& c, const V3<R>& pos, const V3<R>& k,
                  const V3<R>& u, const V3<R>& v) {
  R od[7], lt[7];
  R track = R(c.initial_step), W = R(c.ui_weight);
  od[0] = 0;
  lt[0] = 0;
  size_t i = 0;
  auto acc = [&](int j, const V3<R>& dir, const DosSection& s) {
    R tau = gauss_ext<R>(P, pos + dir * track, R(s.level));
    R at = tau * R(s.amplitude);
    od[j] += (lt[j] + at) * R(s.d_integral) * W;
    lt[j] = at;
  };
  for (int q = 0; q < T.n[0]; ++q, ++i) {
    acc(0, k, T.s[i]);
    track += R(T.s[i].interval);
  }
  if (T.n[1] + T.n[2] == 0) return std::exp(-od[0]);
  od[2] = od[0]; od[1] = od[0];
  lt[2] = lt[0]; lt[1] = lt[0];
  V3<R> d3[3];
  for (int j = 0; j < 3; ++j) d3[j] = k * R(T.axes[j][2]) + u * R(T.axes[j][1]) + v * R(T.axes[j][0]);
  for (int q = 0; q < T.n[1]; ++q, ++i) {
    for (int j = 0; j < 3; ++j) acc(j, d3[j], T.s[i]);
    track += R(T.s[i].interval);
  }
  if (T.n[2] == 0) return (std::exp(-od[0]) + std::exp(-od[1]) + std::exp(-od[2])) / R(3);
  // transform 3 to 7 (:127-137)
  od[6] = od[5] = od[2];
  od[4] = od[3] = od[1];
  R avg = (od[2] + od[1] + od[0]) / R(3);
  od[2] = od[1] = od[0];
  od[0] = avg;
  lt[6] = lt[5] = lt[2];
  lt[4] = lt[3] = lt[1];
  R avgt = (lt[2] + lt[1] + lt[0]) / R(3);
  lt[2] = lt[1] = lt[0];
  lt[0] = avgt;
  V3<R> d7[7];
  for (int j = 0; j < 7; ++j) d7[j] = k * R(T.axes[3 + j][2]) + u * R(T.axes[3 + j][1]) + v * R(T.axes[3 + j][0]);
  for (int q = 0; q < T.n[2]; ++q, ++i) {
    for (int j = 0; j < 7; ++j) acc(j, d7[j], T.s[i]);
    track += R(T.s[i].interval);
  }
  R w7 = R(T.adj7);
  return (std::exp(-od[0]) + (std::exp(-od[1]) + std::exp(-od[2]) + std::exp(-od[3]) + std::exp(-od[4]) +
                              std::exp(-od[5]) + std::exp(-od[6])) * w7) / (R(1) + w7 * R(6));
}

struct DosParamsO {
  int apply_occ = 1, apply_sdw = 0, shadow_type = 0;
  double spot_cos = 0;
  DosConeSpec occ, sdw;
  double ka = 0.5, kd = 0.5;
  double eye[3] = {0, 0, 0};
  double lpos[3] = {0, 0, 0}, lfwd[3] = {0, 0, 1}, lup[3] = {0, 1, 0}, lright[3] = {1, 0, 0};
};

template <class R> struct DosTracer {
  const Pyramid& P;
  const DosParamsO& p;
  DosTableO to, ts;
  V3<R> G;
  bool ok;
  DosTracer(const Pyramid& P_, const DosParamsO& p_, double sigma0) : P(P_), p(p_) {
    ok = dos_build(p.occ, sigma0, to) && dos_build(p.sdw, sigma0, ts);
    G = V3<R>(R(P.G[0]), R(P.G[1]), R(P.G[2]));
  }
  // OcclusionEvaluationKernel (:324-333): k towards the eye, u/v the camera-ray frame.
  R occlusion(const V3<R>& tx, const V3<R>& up, const V3<R>& right) const {
    V3<R> eye(R(p.eye[0]), R(p.eye[1]), R(p.eye[2]));
    V3<R> k = normalize(eye - (tx - G * R(0.5)));
    return dos_cone<R>(P, to, p.occ, tx, k, up, right);
  }
  // ShadowEvaluationKernel (:533-562), with Cone1RayShadow's parameter order (k, v, u).
  R shadow(const V3<R>& tx, Margin* mg) const {
    V3<R> fwd(R(p.lfwd[0]), R(p.lfwd[1]), R(p.lfwd[2]));
    V3<R> up(R(p.lup[0]), R(p.lup[1]), R(p.lup[2])), right(R(p.lright[0]), R(p.lright[1]), R(p.lright[2]));
    if (p.shadow_type == 2) return dos_cone<R>(P, ts, p.sdw, tx, fwd, up, right);
    V3<R> lp(R(p.lpos[0]), R(p.lpos[1]), R(p.lpos[2]));
    V3<R> k = normalize(lp - (tx - G * R(0.5)));
    V3<R> u = normalize(cross(k, right));
    V3<R> v = normalize(cross(k, u));
    if (p.shadow_type == 1) {
      if (mg) mg->see_a(double(dot(k, fwd)), p.spot_cos);
      if (dot(k, fwd) < R(p.spot_cos)) return 0;
    }
    return dos_cone<R>(P, ts, p.sdw, tx, k, v, u);
  }
};

template <class R>
inline void render_dos(const Volume& vol, const TF& tf, const Pyramid& P, const DosParamsO& dp, double sigma0,
                       const CamIn& cam, int W, int H, R step, R et, R* out, double* margin, int stride = 1,
                       const PhongO* ph = nullptr) {
  PhongO Ph = ph ? *ph : PhongO{};
  for (int i = 0; i < 3; ++i) Ph.eye[i] = cam.eye[i];
  DosParamsO q = dp;
  for (int i = 0; i < 3; ++i) q.eye[i] = cam.eye[i];
  DosTracer<R> tr(P, q, sigma0);
  CamBasis<R> cb = cam_basis<R>(cam, W, H);
  V3<R> G(R(vol.nx * vol.sx), R(vol.ny * vol.sy), R(vol.nz * vol.sz));
  R ka = q.apply_occ ? R(q.ka) : R(0), kd = q.apply_sdw ? R(q.kd) : R(0);
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
        V3<R> right = normalize(cross(d, V3<R>(0, 1, 0)));
        V3<R> up = normalize(cross(-d, right));
        auto sh = [&](const V3<R>& p, R* src) {
          R Io = q.apply_occ ? tr.occlusion(p, up, right) : R(0);
          R Is = q.apply_sdw ? tr.shadow(p, &mg) : R(0);
          if (ph) return shade_lit_phong(Ph, G, p, Io, Is, ka, kd, q.apply_sdw ? R(Ph.ks) : R(0), src);  // :627-650
          for (int c = 0; c < 3; ++c) src[c] = (R(1) / (ka + kd)) * (src[c] * Io * ka + src[c] * Is * kd);
        };
        march(tp, d, D, step, et, G, vol, tf.rgbt.data(), tf.size(), sh, o, &mg);
      }
      if (margin) { margin[(size_t(py) * W + px) * 2] = mg.a; margin[(size_t(py) * W + px) * 2 + 1] = mg.g; }
    }
}

// ---------------------------------------------------------------- cone ground truth
// crtgtrenderer.cpp:131-185: theta ~ U[0, aperture], phi ~ U[0, 2 pi) from one
// default-seeded engine (occlusion rays first, then shadow rays); the ray is (0,0,1)
// rotated about y by theta and then about z by phi.
inline void gt_draw(std::default_random_engine& gen, int n, float aperture_deg, std::vector<double>& out) {
  std::uniform_real_distribution<float> U(0.0f, 1.0f);
  const float pi = 3.14159265358979323846f;
  out.resize(size_t(n) * 3);
  for (int i = 0; i < n; ++i) {
    float theta = U(gen) * (pi * (aperture_deg / 180.f));
    float phi = U(gen) * (pi * (360.f / 180.f));
    V3<double> v = rodrigues(V3<double>(0, 0, 1), double(theta), V3<double>(0, 1, 0));
    v = rodrigues(v, double(phi), V3<double>(0, 0, 1));
    out[size_t(i) * 3] = v.x;
    out[size_t(i) * 3 + 1] = v.y;
    out[size_t(i) * 3 + 2] = v.z;
  }
}

struct GtParamsO {
  int apply_occ = 1, apply_sdw = 1, shadow_type = 0;
  std::vector<double> occ_rays, sdw_rays;  // (x, y, z) per ray
  double occ_dist = 0, sdw_dist = 0, gap = 1.0, rstep = 0.5;
  double ka = 0.5, kd = 0.5;
  double lpos[3] = {0, 0, 0}, lfwd[3] = {0, 0, 1}, lup[3] = {0, 1, 0}, lright[3] = {1, 0, 0};  // lfwd as the shader sees it
};

// One ray of ConeOcclusionEvaluationRayCasting / ConeShadowsEvaluationRayCasting
// (gt_ray_marching.comp:125-159, 209-242): trapezoid transmittance over the RGBt
// extinction; stops before a step that would leave the box, or at opacity 0.99.
template <class R>
inline R gt_ray(const Volume& vol, const TF& tf, const V3<R>& G, const V3<R>& tx, const V3<R>& w, R dist, R gap,
                R rstep, Margin* mg) {
  V3<R> invG(R(1) / G.x, R(1) / G.y, R(1) / G.z);
  const float* tab = tf.rgbt.data();
  const int T = tf.size();
  R Vt = 1;
  R s = gap;
  R c[4];
  tex1d_tf<R>(tab, T, tex3d_volume(vol, (tx + w * s) * invG), c);
  R st0 = c[3];
  while (s < dist) {
    R h = std::min(rstep, dist - s);
    V3<R> at = tx + w * (s + h);
    if (mg) {
      const R q[6] = {at.x, at.y, at.z, at.x - G.x, at.y - G.y, at.z - G.z};
      for (R e : q) mg->see_g(double(e), 0.0);
    }
    if (at.x < 0 || at.x > G.x || at.y < 0 || at.y > G.y || at.z < 0 || at.z > G.z) break;
    tex1d_tf<R>(tab, T, tex3d_volume(vol, at * invG), c);
    R st1 = c[3];
    Vt *= std::exp(-((st0 + st1) * R(0.5)) * h);
    if (mg) mg->see_a(double(1 - Vt), 0.99);
    if ((1 - Vt) > R(0.99)) break;
    st0 = st1;
    s = s + h;
  }
  return Vt;
}

template <class R>
inline R gt_cone(const Volume& vol, const TF& tf, const V3<R>& G, const std::vector<double>& rays, const V3<R>& tx,
                 const V3<R>& right, const V3<R>& up, const V3<R>& axis, R dist, R gap, R rstep, Margin* mg) {
  R S = 0, Wt = 0;
  for (size_t r = 0; r < rays.size() / 3; ++r) {
    V3<R> w = normalize(right * R(rays[3 * r]) + up * R(rays[3 * r + 1]) + axis * R(rays[3 * r + 2]));
    R Vt = gt_ray<R>(vol, tf, G, tx, w, dist, gap, rstep, mg);
    R rw = dot(axis, w);
    S += Vt * rw;
    Wt += rw;
  }
  return S / Wt;
}

template <class R>
inline void render_gt(const Volume& vol, const TF& tf, const GtParamsO& q, const CamIn& cam, int W, int H, R step,
                      R et, R* out, double* margin, int stride = 1) {
  CamBasis<R> cb = cam_basis<R>(cam, W, H);
  V3<R> G(R(vol.nx * vol.sx), R(vol.ny * vol.sy), R(vol.nz * vol.sz));
  R ka = q.apply_occ ? R(q.ka) : R(0), kd = q.apply_sdw ? R(q.kd) : R(0);
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
        R D = tf_ - tn;
        V3<R> tp = cb.eye + d * tn + G * R(0.5);
        V3<R> right = normalize(cross(d, V3<R>(0, 1, 0)));
        V3<R> up = normalize(cross(-d, right));
        V3<R> vdir = normalize(-d);
        auto sh = [&](const V3<R>& p, R* src) {
          R Io = 0, Is = 0;
          if (q.apply_occ)
            Io = gt_cone<R>(vol, tf, G, q.occ_rays, p, right, up, vdir, R(q.occ_dist), R(q.gap), R(q.rstep), &mg);
          if (q.apply_sdw && q.shadow_type != 1) {  // the spot branch returns 0 (:192)
            V3<R> sd, su, sr;
            if (q.shadow_type == 2) {
              sd = V3<R>(R(q.lfwd[0]), R(q.lfwd[1]), R(q.lfwd[2]));
              su = V3<R>(R(q.lup[0]), R(q.lup[1]), R(q.lup[2]));
              sr = V3<R>(R(q.lright[0]), R(q.lright[1]), R(q.lright[2]));
            } else {
              V3<R> lp(R(q.lpos[0]), R(q.lpos[1]), R(q.lpos[2]));
              sd = normalize(lp - (p - G * R(0.5)));
              su = normalize(cross(sd, V3<R>(R(q.lright[0]), R(q.lright[1]), R(q.lright[2]))));
              sr = normalize(cross(sd, su));
            }
            Is = gt_cone<R>(vol, tf, G, q.sdw_rays, p, sr, su, sd, R(q.sdw_dist), R(q.gap), R(q.rstep), &mg);
          }
          for (int c = 0; c < 3; ++c) src[c] = (R(1) / (ka + kd)) * (src[c] * Io * ka + src[c] * Is * kd);
        };
        march(tp, d, D, step, et, G, vol, tf.rgbt.data(), tf.size(), sh, o, &mg);
      }
      if (margin) { margin[(size_t(py) * W + px) * 2] = mg.a; margin[(size_t(py) * W + px) * 2 + 1] = mg.g; }
    }
}

}  // namespace orc
