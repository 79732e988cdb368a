// CPU ORACLE — test infrastructure only.
//
// A scalar C++ restatement of the reference's single-pass structured ray
// casters, templated on the arithmetic type `R`:
//   R = double  -> the FP64 correctness oracle the HIP kernels are checked against;
//   R = float   -> the FP32 CPU baseline timed beside the GPU.
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
// this code.  The product library never links it.
//
// Every function cites the reference file:line it follows (paths relative to
// the reference repository root).  GL texture sampling (GL_LINEAR,
// CLAMP_TO_EDGE, LINEAR_MIPMAP_LINEAR) is emulated with exact arithmetic in R;
// the reference's GPU storage formats (R16F volume, RGBA16F tables) are not
// emulated: textures hold the float values the reference's C++ code uploads.
//
// Discrete decisions (early ray termination, EBS volume-boundary cut, EBS
// ceil()) are ill-conditioned when the compared quantities are nearly equal.
// Each render records, per pixel, the smallest |lhs - rhs| seen at such a
// decision (the pixel's "margin"); the parity tests treat pixels whose FP64
// margin lies below the FP32 resolution of that quantity as straddles.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>

namespace orc {

// ---------------------------------------------------------------- vectors
template <class R> struct V3 {
  R x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
  template <class S> explicit V3(const V3<S>& o) : x(R(o.x)), y(R(o.y)), z(R(o.z)) {}
  V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
  V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
  V3 operator*(const V3& o) const { return V3(x * o.x, y * o.y, z * o.z); }
  V3 operator/(const V3& o) const { return V3(x / o.x, y / o.y, z / o.z); }
  V3 operator*(R s) const { return V3(x * s, y * s, z * s); }
  V3 operator-() const { return V3(-x, -y, -z); }
};
template <class R> inline R dot(const V3<R>& a, const V3<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(const V3<R>& a, const V3<R>& b) {
  return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> inline V3<R> normalize(const V3<R>& v) {
  R l = std::sqrt(dot(v, v));
  return V3<R>(v.x / l, v.y / l, v.z / l);
}
template <class R> inline V3<R> vabs(const V3<R>& v) { return V3<R>(std::fabs(v.x), std::fabs(v.y), std::fabs(v.z)); }
template <class R> inline V3<R> vmin(const V3<R>& a, const V3<R>& b) {
  return V3<R>(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z));
}
template <class R> inline V3<R> vmax(const V3<R>& a, const V3<R>& b) {
  return V3<R>(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z));
}
template <class R> inline V3<R> vclamp(const V3<R>& v, const V3<R>& lo, const V3<R>& hi) { return vmin(vmax(v, lo), hi); }

// Per-pixel record of the closest call at a discrete decision.
// `a`: opacity / transmittance thresholds (dimensionless); `g`: geometric tests (world units).
struct Margin {
  double a = std::numeric_limits<double>::infinity();
  double g = std::numeric_limits<double>::infinity();
  inline void see_a(double x, double y) { a = std::min(a, std::fabs(x - y)); }
  inline void see_g(double x, double y) { g = std::min(g, std::fabs(x - y)); }
};

// ---------------------------------------------------------------- volume
// libs/volvis_utils/structuredgridvolume.cpp:121-151 (GetNormalizedSample):
// x-fastest index, v/255 (8 bit) or v/65535 (16 bit), 0 outside the grid.
struct Volume {
  int nx = 0, ny = 0, nz = 0;
  int dtype = 0;  // 0 = uint8, 1 = uint16
  const void* data = nullptr;
  double sx = 1, sy = 1, sz = 1;
  template <class R> inline R norm(int x, int y, int z) const {
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return R(0);
    size_t i = size_t(x) + size_t(y) * nx + size_t(z) * size_t(nx) * ny;
    if (dtype == 0) return R(static_cast<const uint8_t*>(data)[i]) / R(255.0);
    return R(static_cast<const uint16_t*>(data)[i]) / R(65535.0);
  }
};

// GL_LINEAR + CLAMP_TO_EDGE along one axis (OpenGL 4.3 spec, 8.14.2):
// texel coordinate u*n - 1/2, floor/frac, both neighbours clamped to [0, n-1].
template <class R> inline void gl_linear_axis(R u, int n, int& i0, int& i1, R& f) {
  R x = u * R(n) - R(0.5);
  R fl = std::floor(x);
  f = x - fl;
  long long i = (long long)fl;
  long long j = i + 1;
  i0 = int(std::min<long long>(std::max<long long>(i, 0), n - 1));
  i1 = int(std::min<long long>(std::max<long long>(j, 0), n - 1));
}

// 3D GL_LINEAR texture of the normalised volume (libs/volvis_utils/utils.cpp:20-56,
// GL_LINEAR / CLAMP_TO_EDGE at :8-9).
template <class R> inline R tex3d_volume(const Volume& v, const V3<R>& uvw) {
  int x0, x1, y0, y1, z0, z1;
  R fx, fy, fz;
  gl_linear_axis(uvw.x, v.nx, x0, x1, fx);
  gl_linear_axis(uvw.y, v.ny, y0, y1, fy);
  gl_linear_axis(uvw.z, v.nz, z0, z1, fz);
  R c000 = v.norm<R>(x0, y0, z0), c100 = v.norm<R>(x1, y0, z0);
  R c010 = v.norm<R>(x0, y1, z0), c110 = v.norm<R>(x1, y1, z0);
  R c001 = v.norm<R>(x0, y0, z1), c101 = v.norm<R>(x1, y0, z1);
  R c011 = v.norm<R>(x0, y1, z1), c111 = v.norm<R>(x1, y1, z1);
  R c00 = c000 * (1 - fx) + c100 * fx, c10 = c010 * (1 - fx) + c110 * fx;
  R c01 = c001 * (1 - fx) + c101 * fx, c11 = c011 * (1 - fx) + c111 * fx;
  R c0 = c00 * (1 - fy) + c10 * fy, c1 = c01 * (1 - fy) + c11 * fy;
  return c0 * (1 - fz) + c1 * fz;
}

// 3D GL_LINEAR texture over a dense scalar grid (used for the SAT, ebsrenderer.cpp:715-717,
// and the mip levels of the extinction pyramid).
template <class R, class T> inline R tex3d_grid(const T* g, int nx, int ny, int nz, const V3<R>& uvw) {
  int x0, x1, y0, y1, z0, z1;
  R fx, fy, fz;
  gl_linear_axis(uvw.x, nx, x0, x1, fx);
  gl_linear_axis(uvw.y, ny, y0, y1, fy);
  gl_linear_axis(uvw.z, nz, z0, z1, fz);
  auto at = [&](int x, int y, int z) { return R(g[size_t(x) + size_t(y) * nx + size_t(z) * size_t(nx) * ny]); };
  R c00 = at(x0, y0, z0) * (1 - fx) + at(x1, y0, z0) * fx;
  R c10 = at(x0, y1, z0) * (1 - fx) + at(x1, y1, z0) * fx;
  R c01 = at(x0, y0, z1) * (1 - fx) + at(x1, y0, z1) * fx;
  R c11 = at(x0, y1, z1) * (1 - fx) + at(x1, y1, z1) * fx;
  R c0 = c00 * (1 - fy) + c10 * fy, c1 = c01 * (1 - fy) + c11 * fy;
  return c0 * (1 - fz) + c1 * fz;
}

// ---------------------------------------------------------------- transfer function
// libs/volvis_utils/transferfunction1d.cpp.  Control point colours are stored in a
// float glm::vec4 (transferfunction.cpp:8-24), the table is glm::dvec4 zero-initialised.
struct TF {
  int max_density = 255;
  bool ext_input = false;
  std::vector<double> tab;   // (max_density+1) x 4, BuildLinear (transferfunction1d.cpp:319-358)
  std::vector<float> rgbt;   // GenerateTexture_1D_RGBt upload array (transferfunction1d.cpp:89-118)
  std::vector<float> rgba;   // GenerateTexture_1D_RGBA upload array (transferfunction1d.cpp:58-87)
  int size() const { return max_density + 1; }
};

// transferfunction.h:74-82
inline double ext_to_opacity(float e) { return 1.0 - std::exp(-double(e)); }
inline double opacity_to_ext(float o) { return std::log(1.0 / (1.0 - double(o))); }

// Returns false on control points the reference would turn into NaN entries
// (two consecutive points with the same isovalue) or that index outside the table.
inline bool tf_build(TF& tf, int max_density, bool ext_input, int n_rgb, const double* rgb_iso, int n_a,
                     const double* a_iso) {
  tf.max_density = max_density;
  tf.ext_input = ext_input;
  int T = max_density + 1;
  tf.tab.assign(size_t(T) * 4, 0.0);
  // BuildLinear, rgb part (:321-340)
  for (int i = 0; i + 1 < n_rgb; ++i) {
    int i0 = int(rgb_iso[i * 4 + 3]), i1 = int(rgb_iso[(i + 1) * 4 + 3]);
    if (i1 == i0 || i0 < 0 || i1 >= T || i0 >= T || i1 < 0) return false;
    float c0[3] = {float(rgb_iso[i * 4 + 0]), float(rgb_iso[i * 4 + 1]), float(rgb_iso[i * 4 + 2])};
    float c1[3] = {float(rgb_iso[(i + 1) * 4 + 0]), float(rgb_iso[(i + 1) * 4 + 1]), float(rgb_iso[(i + 1) * 4 + 2])};
    double diff[3] = {double(c1[0] - c0[0]), double(c1[1] - c0[1]), double(c1[2] - c0[2])};
    for (int x = i0; x <= i1; ++x) {
      double k = double(x - i0) / double(i1 - i0);
      for (int c = 0; c < 3; ++c) tf.tab[size_t(x) * 4 + c] = double(c0[c]) + diff[c] * k;
    }
  }
  // alpha part (:342-357)
  for (int i = 0; i + 1 < n_a; ++i) {
    int i0 = int(a_iso[i * 2 + 1]), i1 = int(a_iso[(i + 1) * 2 + 1]);
    if (i1 == i0 || i0 < 0 || i1 >= T || i0 >= T || i1 < 0) return false;
    float a0 = float(a_iso[i * 2 + 0]), a1 = float(a_iso[(i + 1) * 2 + 0]);
    double diff = double(a1 - a0);
    for (int x = i0; x <= i1; ++x) {
      double k = double(x - i0) / double(i1 - i0);
      tf.tab[size_t(x) * 4 + 3] = double(a0) + diff * k;
    }
  }
  tf.rgbt.resize(size_t(T) * 4);
  tf.rgba.resize(size_t(T) * 4);
  for (int i = 0; i < T; ++i) {
    for (int c = 0; c < 3; ++c) {
      tf.rgbt[i * 4 + c] = float(tf.tab[i * 4 + c]);
      tf.rgba[i * 4 + c] = float(tf.tab[i * 4 + c]);
    }
    float v4 = float(tf.tab[i * 4 + 3]);
    tf.rgbt[i * 4 + 3] = ext_input ? v4 : float(opacity_to_ext(v4));
    float w4 = float(tf.tab[i * 4 + 3]);
    tf.rgba[i * 4 + 3] = ext_input ? float(ext_to_opacity(w4)) : w4;
  }
  return true;
}

// TransferFunction1D::Get (transferfunction1d.cpp:132-157), returned as float vec4.
inline void tf_get(const TF& tf, double value, double max_data_value, float out[4]) {
  if (max_data_value >= 0) value = value * (double(tf.max_density) / max_data_value);
  if (value < 0.0f || value > float(tf.max_density)) {
    out[0] = out[1] = out[2] = out[3] = 0.f;
    return;
  }
  if (std::fabs(value - float(tf.max_density)) < 0.000001) {
    for (int c = 0; c < 4; ++c) out[c] = float(tf.tab[size_t(tf.max_density) * 4 + c]);
    return;
  }
  int i = int(value);
  double t = value - i;
  for (int c = 0; c < 4; ++c) out[c] = float((1.0 - t) * tf.tab[size_t(i) * 4 + c] + t * tf.tab[size_t(i + 1) * 4 + c]);
}

// GetExtN (transferfunction1d.cpp:189-197)
inline float tf_get_ext_n(const TF& tf, double normalized_value) {
  float v[4];
  tf_get(tf, normalized_value, 1.0, v);
  float val = v[3];
  if (!tf.ext_input) return float(opacity_to_ext(val));
  return val;
}

// texture(TexTransferFunc, d) over an RGBA 1D table with GL_LINEAR/CLAMP_TO_EDGE
// (texture1d.cpp:19-28).
template <class R> inline void tex1d_tf(const float* table, int T, R d, R out[4]) {
  int i0, i1;
  R f;
  gl_linear_axis(d, T, i0, i1, f);
  for (int c = 0; c < 4; ++c) out[c] = R(table[i0 * 4 + c]) * (1 - f) + R(table[i1 * 4 + c]) * f;
}

// ---------------------------------------------------------------- camera / rays
struct CamIn {
  float eye[3], center[3], up[3];
  float fovy_deg, aspect;
};

// glm::lookAt basis (vis_utils/camera.cpp:281-284; glm 0.9.5 right-handed):
// f = normalize(center-eye), s = normalize(cross(f, up)), u = cross(s, f).
template <class R> struct CamBasis {
  V3<R> eye, s, u, f;
  R tanfov, aspect;
};
template <class R> inline CamBasis<R> cam_basis(const CamIn& c, int W, int H) {
  CamBasis<R> b;
  b.eye = V3<R>(c.eye[0], c.eye[1], c.eye[2]);
  V3<R> ce(c.center[0], c.center[1], c.center[2]), up(c.up[0], c.up[1], c.up[2]);
  b.f = normalize(ce - b.eye);
  b.s = normalize(cross(b.f, up));
  b.u = cross(b.s, b.f);
  // uniform u_TanCameraFovY is a float (rc1prenderer.cpp:97); aspect is W/H in float (camera.cpp:307-310)
  b.tanfov = R(float(std::tan((double(c.fovy_deg) * (3.14159265358979323846 / 180.0)) / 2.0)));
  b.aspect = c.aspect > 0 ? R(c.aspect) : R(float(W) / float(H));
  return b;
}

// ray_marching_1p.comp:158-170: camera_dir = normalize(vec3(Vx*tan*aspect, Vy*tan, -1) * mat3(View)).
// The row-vector product with glm's view matrix expands to a*s + b*u + f.
template <class R> inline V3<R> ray_dir(const CamBasis<R>& b, int px, int py, int W, int H) {
  R fx = R(px) + R(0.5), fy = R(py) + R(0.5);
  R vx = (fx / R(W)) * 2 - 1, vy = (fy / R(H)) * 2 - 1;
  R a = vx * b.tanfov * b.aspect, c = vy * b.tanfov;
  return normalize(b.s * a + b.u * c + b.f);
}

// _common_shaders/ray_bbox_intersection.comp:18-52 (slab test, then tnear = max(tnear, 0)).
template <class R> inline bool ray_aabb(const V3<R>& o, const V3<R>& d, const V3<R>& G, R& tnear, R& tfar) {
  V3<R> bmin = G * R(-0.5), bmax = G * R(0.5);
  V3<R> inv(R(1) / d.x, R(1) / d.y, R(1) / d.z);
  V3<R> t0 = inv * (bmin - o), t1 = inv * (bmax - o);
  V3<R> tmin = vmin(t0, t1), tmax = vmax(t0, t1);
  tnear = std::max(std::max(tmin.x, tmin.y), tmin.z);
  tfar = std::min(std::min(tmax.x, tmax.y), tmax.z);
  bool hit = tfar > tnear;
  tnear = std::max(tnear, R(0));
  return hit;
}

// ---------------------------------------------------------------- plain compositing
// One front-to-back march, ray_marching_1p.comp:177-243 (USE_TRANSPARENCY undefined).
// `shade` may rewrite src.rgb for opaque samples (the EBS/cone variants).
template <class R, class Shade>
inline void march(const V3<R>& tex_pos, const V3<R>& dir, R D, R step, R et, const V3<R>& G, const Volume& vol,
                  const float* tf, int T, Shade&& shade, R dst[4], Margin* mg, int mid_sample = 1) {
  dst[0] = dst[1] = dst[2] = dst[3] = 0;
  V3<R> invG(R(1) / G.x, R(1) / G.y, R(1) / G.z);
  for (R s = 0; s < D;) {
    R h = std::min(step, D - s);
    V3<R> p = tex_pos + dir * (mid_sample ? (s + h * R(0.5)) : (s + h));
    R density = tex3d_volume(vol, p * invG);
    R src[4];
    tex1d_tf<R>(tf, T, density, src);
    if (src[3] > 0) {
      shade(p, src);
      R a = 1 - std::exp(-src[3] * h);
      R w = 1 - dst[3];
      dst[0] += w * src[0] * a;
      dst[1] += w * src[1] * a;
      dst[2] += w * src[2] * a;
      dst[3] += w * a;
      if (mg) mg->see_a(double(dst[3]), double(et));
      if (dst[3] > et) break;
    }
    s = s + h;
  }
}

// ---------------------------------------------------------------- Blinn-Phong
// Gradient-based shading: ShadeBlinnPhong (rc1pass/ray_marching_1p.comp:48-81) and the
// Phong branches of ebs_ray_bbox_marching.comp:523-546, rc1pdosct/ray_bbox_marching.comp:
// 627-650 and _common_shaders/obj_ray_marching.comp:186-211.  The gradient texture is
// sampled with GL_LINEAR / CLAMP_TO_EDGE at tex_pos / G; a zero vector means the sample
// keeps its colour.
struct GradientO {
  int nx = 0, ny = 0, nz = 0;
  std::vector<double> gx, gy, gz;  // x fastest
};

struct PhongO {
  const GradientO* g = nullptr;
  double ka = 0.5, kd = 0.5, ks = 0.8, shininess = 30;  // renderingparameters.cpp:23-26
  double spec[3] = {1, 1, 1};                            // lightsourcelist.cpp:24
  double lpos[3] = {0, 0, 0};
  double eye[3] = {0, 0, 0};
};

// max(0, n.l) and max(0, h.n)^shininess at texture-space p; false where the GL_LINEAR
// gradient is the zero vector.
template <class R>
inline bool phong_terms(const PhongO& P, const V3<R>& G, const V3<R>& p, R& dd, R& sp) {
  const V3<R> uvw(p.x / G.x, p.y / G.y, p.z / G.z);
  V3<R> n(tex3d_grid<R, double>(P.g->gx.data(), P.g->nx, P.g->ny, P.g->nz, uvw),
          tex3d_grid<R, double>(P.g->gy.data(), P.g->nx, P.g->ny, P.g->nz, uvw),
          tex3d_grid<R, double>(P.g->gz.data(), P.g->nx, P.g->ny, P.g->nz, uvw));
  if (n.x == 0 && n.y == 0 && n.z == 0) return false;
  n = normalize(n);
  const V3<R> wp = p - G * R(0.5);
  const V3<R> L = normalize(V3<R>(R(P.lpos[0]), R(P.lpos[1]), R(P.lpos[2])) - wp);
  const V3<R> E = normalize(V3<R>(R(P.eye[0]), R(P.eye[1]), R(P.eye[2])) - wp);
  const V3<R> H = normalize(E + L);
  dd = std::max(R(0), dot(n, L));
  sp = std::pow(std::max(R(0), dot(H, n)), R(P.shininess));
  return true;
}

// ShadeBlinnPhong (ray_marching_1p.comp:48-81)
template <class R> inline void shade_blinn_phong(const PhongO& P, const V3<R>& G, const V3<R>& p, R* rgb) {
  R dd, sp;
  if (!phong_terms(P, G, p, dd, sp)) return;
  for (int c = 0; c < 3; ++c) rgb[c] = rgb[c] * (R(P.ka) + R(P.kd) * dd) + R(P.spec[c]) * R(P.ks) * sp;
}

template <class R>
inline void shade_lit_phong(const PhongO& P, const V3<R>& G, const V3<R>& p, R Io, R Is, R ka, R kd, R ks, R* rgb) {
  R dd, sp;
  if (!phong_terms(P, G, p, dd, sp)) return;
  for (int c = 0; c < 3; ++c)
    rgb[c] = (R(1) / (ka + kd)) * (rgb[c] * Io * ka + Is * (rgb[c] * kd * dd)) + Is * (ks * R(P.spec[c]) * sp);
}

// The Phong branch of a lit marcher: occlusion Io, shadow Is, effective weights ka, kd,
// ks (0 for a term that is off); the DOS shader writes the same quantity as
// rgb*((Io*ka + Is*kd*dd)/(ka+kd)) + spec*Is*ks*sp.
template <class R>
inline void render_rc1p(const Volume& vol, const TF& tf, const CamIn& cam, int W, int H, R step, R et, R* out,
                        double* margin, int stride = 1) {
  CamBasis<R> cb = cam_basis<R>(cam, W, H);
  V3<R> G(R(vol.nx * vol.sx), R(vol.ny * vol.sy), R(vol.nz * vol.sz));
  auto noshade = [](const V3<R>&, R*) {};
#pragma omp parallel for schedule(static)
  for (int py = 0; py < H; ++py)
    for (int px = 0; px < W; ++px) {
      R* o = out + (size_t(py) * W + px) * 4;
      o[0] = o[1] = o[2] = o[3] = 0;
      if (px % stride || py % stride) continue;  // sparse subset of pixels
      Margin mg;
      V3<R> d = ray_dir(cb, px, py, W, H);
      R tn, tf_;
      if (ray_aabb(cb.eye, d, G, tn, tf_)) {
        R D = std::fabs(tf_ - tn);
        V3<R> tp = cb.eye + d * tn + G * R(0.5);
        march(tp, d, D, step, et, G, vol, tf.rgbt.data(), tf.size(), noshade, o, &mg);
      }
      if (margin) { margin[(size_t(py) * W + px) * 2] = mg.a; margin[(size_t(py) * W + px) * 2 + 1] = mg.g; }
    }
}

// ---------------------------------------------------------------- summed-area table
// ebsrenderer.cpp:624-663: (W+2)(H+2)(D+2) grid, zero border, interior GetExtN(sample(x-1,y-1,z-1)).
inline void sat_input(const Volume& vol, const TF& tf, std::vector<double>& g) {
  int w = vol.nx + 2, h = vol.ny + 2, d = vol.nz + 2;
  g.assign(size_t(w) * h * d, 0.0);
#pragma omp parallel for schedule(static)
  for (int z = 1; z < d - 1; ++z)
    for (int y = 1; y < h - 1; ++y)
      for (int x = 1; x < w - 1; ++x)
        g[size_t(x) + size_t(y) * w + size_t(z) * w * h] =
            double(tf_get_ext_n(tf, vol.norm<double>(x - 1, y - 1, z - 1)));
}

// SummedAreaTable3D<double>::BuildSAT (vis_utils/summedareatable.h:218-278), the same
// four-step recurrence with the clamped GetValue (:204-213).
inline void sat_build_recurrence(std::vector<double>& g, int w, int h, int d) {
  auto at = [&](int x, int y, int z) -> double {
    if (x < 0 || y < 0 || z < 0) return 0.0;
    if (x >= w) x = w - 1;
    if (y >= h) y = h - 1;
    if (z >= d) z = d - 1;
    return g[size_t(x) + size_t(w) * y + size_t(w) * h * z];
  };
  auto set = [&](double v, int x, int y, int z) { g[size_t(x) + size_t(w) * y + size_t(w) * h * z] = v; };
  for (int x = 1; x < w; x++) set(at(x - 1, 0, 0) + at(x, 0, 0), x, 0, 0);
  for (int y = 1; y < h; y++) set(at(0, y - 1, 0) + at(0, y, 0), 0, y, 0);
  for (int z = 1; z < d; z++) set(at(0, 0, z - 1) + at(0, 0, z), 0, 0, z);
  for (int x = 1; x < w; x++)
    for (int z = 1; z < d; z++) set(at(x - 1, 0, z) + at(x, 0, z - 1) - at(x - 1, 0, z - 1) + at(x, 0, z), x, 0, z);
  for (int x = 1; x < w; x++)
    for (int y = 1; y < h; y++) set(at(x - 1, y, 0) + at(x, y - 1, 0) - at(x - 1, y - 1, 0) + at(x, y, 0), x, y, 0);
  for (int y = 1; y < h; y++)
    for (int z = 1; z < d; z++) set(at(0, y - 1, z) + at(0, y, z - 1) - at(0, y - 1, z - 1) + at(0, y, z), 0, y, z);
  for (int x = 1; x < w; x++)
    for (int y = 1; y < h; y++)
      for (int z = 1; z < d; z++) {
        double v = at(x, y, z) + at(x - 1, y - 1, z - 1) + at(x, y, z - 1) + at(x, y - 1, z) + at(x - 1, y, z) -
                   at(x - 1, y - 1, z) - at(x, y - 1, z - 1) - at(x - 1, y, z - 1);
        set(v, x, y, z);
      }
}

// ---------------------------------------------------------------- extinction-based shading
// rc1pextbsd/ebs_ray_bbox_marching.comp, constants :65-73, uniforms from ebsrenderer.cpp:125-247.
struct EbsParams {
  int apply_occlusion = 1, apply_shadow = 1;
  int ao_shells = 15;          // ebsrenderer.cpp:30
  float ao_radius = 1.0f;      // :31
  float sdw_angle_rad = 0;     // DirSdwConeAngle, (float)(deg*pi/180) :161
  float sdw_interval = 2.0f;   // :38
  float sdw_initial = 2.0f;    // :39
  float sdw_weight = 1.0f;     // :40
  float sdw_max_dist = 0;      // 0.75*diagonal, :99-105
  int shadow_type = 1;         // 0 point light, 1 directional (LightCamForward)
  float light_pos[3] = {0, 0, 0};
  float light_fwd[3] = {0, 0, 1};
  float ka = 0.5f, kd = 0.5f;  // renderingparameters.cpp:23-24
};

template <class R, class S> struct Ebs {
  const S* sat;  // (nx+2)(ny+2)(nz+2)
  int w, h, d;
  V3<R> scale, G, minSAT, maxSAT, minVol, maxVol, invSat;
  EbsParams p;
  Ebs(const S* sat_, const Volume& v, const EbsParams& prm) : sat(sat_), p(prm) {
    w = v.nx + 2; h = v.ny + 2; d = v.nz + 2;
    scale = V3<R>(R(float(v.sx)), R(float(v.sy)), R(float(v.sz)));
    G = V3<R>(R(v.nx) * scale.x, R(v.ny) * scale.y, R(v.nz) * scale.z);
    minSAT = scale * R(0.5);
    maxSAT = G + scale * R(1.5);
    minVol = scale * R(0.5);
    maxVol = G - scale * R(0.5);
    V3<R> t = G + scale * R(2);
    invSat = V3<R>(R(1) / t.x, R(1) / t.y, R(1) / t.z);
  }
  // :74-83
  R summed(R x, R y, R z) const { return tex3d_grid<R, S>(sat, w, h, d, V3<R>(x, y, z) * invSat); }
  // :87-100
  R eval_sat(const V3<R>& p1, const V3<R>& p2) const {
    R V1 = summed(p2.x, p2.y, p2.z), V2 = summed(p1.x, p2.y, p2.z);
    R V3_ = summed(p2.x, p2.y, p1.z), V4 = summed(p1.x, p2.y, p1.z);
    R V5 = summed(p2.x, p1.y, p2.z), V6 = summed(p1.x, p1.y, p2.z);
    R V7 = summed(p2.x, p1.y, p1.z), V8 = summed(p1.x, p1.y, p1.z);
    return V1 - V2 - V3_ + V4 - V5 + V6 + V7 - V8;
  }
  // :102-110
  R ao_box(V3<R> p1, V3<R> p2) const {
    p1 = vclamp(p1 + scale, minSAT, maxSAT);
    p2 = vclamp(p2 + scale, minSAT, maxSAT);
    return eval_sat(p1, p2);
  }
  // :112-146
  R ambient_occlusion(const V3<R>& tx) const {
    R rad = R(p.ao_radius);
    R S0 = ao_box(tx - scale * rad, tx + scale * rad);
    R t = S0 * (R(1) / (rad * rad));
    R Si = S0;
    for (int i = 1; i < p.ao_shells; ++i) {
      R r1 = rad * R(i + 1);
      R S1 = ao_box(tx - scale * r1, tx + scale * r1);
      t = t + (S1 - Si) * (R(1) / (r1 * r1));
      Si = S1;
    }
    R rn = rad * R(p.ao_shells);
    return std::exp(-((R(1) / (rn * rn)) * t));
  }
  // :148-188 (texture path)
  R shadow_box(V3<R> p1, V3<R> p2) const {
    R vq = (std::fabs(p1.x - p2.x) / scale.x) * (std::fabs(p1.y - p2.y) / scale.y) * (std::fabs(p1.z - p2.z) / scale.z);
    p1 = vclamp(p1 + scale, minSAT, maxSAT);
    p2 = vclamp(p2 + scale, minSAT, maxSAT);
    return (eval_sat(p1, p2) / vq) * R(p.sdw_weight);
  }
  // Generic form of ConeZAxis / ConeYAxis / ConeXAxis (:190-430): axis `a` is the dominant
  // component, (b, c) the secondary axes in the order the reference builds its p1/p2 boxes.
  // The reference's three functions rotate the projected vectors with the same +-angle
  // formulas written per plane; the per-axis helpers below keep each one literal.
  R cone_axis(const V3<R>& pos, const V3<R>& cv, int axis, Margin* mg) const {
    R ang = R(p.sdw_angle_rad);
    R pc = std::cos(ang), ps = std::sin(ang), nc = std::cos(-ang), ns = std::sin(-ang);
    auto comp = [](const V3<R>& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); };
    R Stau = 0;
    R sgn = comp(cv, axis) < 0 ? R(-1) : R(1);
    R sc_a = comp(scale, axis);
    R interval = R(p.sdw_interval) * sgn * sc_a;
    R apos = R(p.sdw_initial) * sgn * sc_a;
    // secondary-plane extents at the mean position: (lo, hi) of the two rotated projections
    // for each secondary axis, as the literal per-axis code computes them.
    V3<R> pj[4];  // pj1, pj2 for first secondary, pj1, pj2 for second secondary
    int sa, sb;   // secondary axes in reference order
    if (axis == 2) {  // ConeZAxis :190-270 — secondaries x (from proj_x) and y (from proj_y)
      sa = 0; sb = 1;
      V3<R> proj_y = normalize(V3<R>(0, cv.y, cv.z));
      V3<R> proj_x = normalize(V3<R>(cv.x, 0, cv.z));
      pj[0] = normalize(V3<R>(proj_x.x * nc - proj_x.z * ns, 0, proj_x.x * ns + proj_x.z * nc));
      pj[1] = normalize(V3<R>(proj_x.x * pc - proj_x.z * ps, 0, proj_x.x * ps + proj_x.z * pc));
      pj[2] = normalize(V3<R>(0, proj_y.y * nc - proj_y.z * ns, proj_y.y * ns + proj_y.z * nc));
      pj[3] = normalize(V3<R>(0, proj_y.y * pc - proj_y.z * ps, proj_y.y * ps + proj_y.z * pc));
    } else if (axis == 1) {  // ConeYAxis :272-350 — secondaries x (proj_x) and z (proj_z)
      sa = 0; sb = 2;
      V3<R> proj_x = normalize(V3<R>(cv.x, cv.y, 0));
      V3<R> proj_z = normalize(V3<R>(0, cv.y, cv.z));
      pj[0] = normalize(V3<R>(proj_x.x * nc - proj_x.y * ns, proj_x.x * ns + proj_x.y * nc, 0));
      pj[1] = normalize(V3<R>(proj_x.x * pc - proj_x.y * ps, proj_x.x * ps + proj_x.y * pc, 0));
      pj[2] = normalize(V3<R>(0, proj_z.z * ns + proj_z.y * nc, proj_z.z * nc - proj_z.y * ns));
      pj[3] = normalize(V3<R>(0, proj_z.z * ps + proj_z.y * pc, proj_z.z * pc - proj_z.y * ps));
    } else {  // ConeXAxis :352-430 — secondaries y (proj_y) and z (proj_z)
      sa = 1; sb = 2;
      V3<R> proj_y = normalize(V3<R>(cv.x, cv.y, 0));
      V3<R> proj_z = normalize(V3<R>(cv.x, 0, cv.z));
      pj[0] = normalize(V3<R>(proj_y.y * ns + proj_y.x * nc, proj_y.y * nc - proj_y.x * ns, 0));
      pj[1] = normalize(V3<R>(proj_y.y * ps + proj_y.x * pc, proj_y.y * pc - proj_y.x * ps, 0));
      pj[2] = normalize(V3<R>(proj_z.z * ns + proj_z.x * nc, 0, proj_z.z * nc - proj_z.x * ns));
      pj[3] = normalize(V3<R>(proj_z.z * ps + proj_z.x * pc, 0, proj_z.z * pc - proj_z.x * ps));
    }
    R pos_a = comp(pos, axis), lo_a = comp(minVol, axis), hi_a = comp(maxVol, axis);
    R sc_sa = comp(scale, sa), sc_sb = comp(scale, sb);
    while (true) {
      R dist = apos / comp(cv, axis);
      if (mg) mg->see_g(double(dist), double(p.sdw_max_dist));
      if (!(dist < R(p.sdw_max_dist))) break;
      R nxt = pos_a + (apos + interval);
      if (mg) { mg->see_g(double(nxt), double(lo_a)); mg->see_g(double(nxt), double(hi_a)); }
      if (!(nxt > lo_a && nxt < hi_a)) break;  // CUT_WHEN_AWAY_FROM_VOLUME
      R mean = std::fabs(apos + interval * R(0.5));
      R pa1 = comp(pj[0], sa) * (mean / std::fabs(comp(pj[0], axis)));
      R pa2 = comp(pj[1], sa) * (mean / std::fabs(comp(pj[1], axis)));
      R pb1 = comp(pj[2], sb) * (mean / std::fabs(comp(pj[2], axis)));
      R pb2 = comp(pj[3], sb) * (mean / std::fabs(comp(pj[3], axis)));
      R a1 = std::min(pa1, pa2), a2 = std::max(pa1, pa2);
      R b1 = std::min(pb1, pb2), b2 = std::max(pb1, pb2);
      R adiff = std::fabs(a2 - a1), bdiff = std::fabs(b2 - b1);
      R qa = adiff / sc_sa, qb = bdiff / sc_sb;
      if (mg) {
        mg->see_g(double(qa), std::round(double(qa)));
        mg->see_g(double(qb), std::round(double(qb)));
      }
      R as = (std::ceil(qa) - qa) * R(0.5);
      R bs = (std::ceil(qb) - qb) * R(0.5);
      a1 = a1 - as * sc_sa; a2 = a2 + as * sc_sa;
      b1 = b1 - bs * sc_sb; b2 = b2 + bs * sc_sb;
      R c1 = std::min(apos, apos + interval), c2 = std::max(apos, apos + interval);
      V3<R> p1, p2;
      auto setc = [](V3<R>& v, int i, R x) { if (i == 0) v.x = x; else if (i == 1) v.y = x; else v.z = x; };
      setc(p1, axis, c1); setc(p2, axis, c2);
      setc(p1, sa, a1); setc(p2, sa, a2);
      setc(p1, sb, b1); setc(p2, sb, b2);
      Stau += shadow_box(pos + p1, pos + p2);
      apos = apos + interval;
    }
    return Stau;
  }
  // :432-456
  R shadows(const V3<R>& tx, Margin* mg) const {
    V3<R> realpos = tx - G * R(0.5);
    V3<R> cv;
    if (p.shadow_type == 0)
      cv = normalize(V3<R>(R(p.light_pos[0]), R(p.light_pos[1]), R(p.light_pos[2])) - realpos);
    else
      cv = normalize(V3<R>(R(p.light_fwd[0]), R(p.light_fwd[1]), R(p.light_fwd[2])));
    V3<R> ac = vabs(cv);
    R st;
    if (ac.z > ac.x && ac.z > ac.y) st = cone_axis(tx, cv, 2, mg);
    else if (ac.y > ac.x) st = cone_axis(tx, cv, 1, mg);
    else st = cone_axis(tx, cv, 0, mg);
    return std::exp(-st);
  }
  // :500-551 (non-Phong branch)
  void shade(const V3<R>& tx, R* src, Margin* mg) const {
    R ka = 0, kd = 0, Io = 0, Is = 0;
    if (p.apply_occlusion) { ka = R(p.ka); Io = ambient_occlusion(tx); }
    if (p.apply_shadow) { kd = R(p.kd); Is = shadows(tx, mg); }
    for (int c = 0; c < 3; ++c) src[c] = (R(1) / (ka + kd)) * (src[c] * Io * ka + src[c] * Is * kd);
  }
};

// ebs_ray_bbox_marching.comp:553-627
template <class R, class S>
inline void render_ebs(const Volume& vol, const TF& tf, const S* sat, const EbsParams& prm, const CamIn& cam, int W,
                       int H, R step, R et, R* out, double* margin, int stride = 1, const PhongO* ph = nullptr) {
  Ebs<R, S> ebs(sat, vol, prm);
  CamBasis<R> cb = cam_basis<R>(cam, W, H);
  V3<R> G = ebs.G;
  PhongO P = ph ? *ph : PhongO{};
  for (int i = 0; i < 3; ++i) P.eye[i] = cam.eye[i];
#pragma omp parallel for schedule(dynamic, 1)
  for (int py = 0; py < H; ++py)
    for (int px = 0; px < W; ++px) {
      R* o = out + (size_t(py) * W + px) * 4;
      o[0] = o[1] = o[2] = o[3] = 0;
      if (px % stride || py % stride) continue;  // sparse subset of pixels
      Margin mg;
      V3<R> d = ray_dir(cb, px, py, W, H);
      R tn, tf_;
      if (ray_aabb(cb.eye, d, G, tn, tf_)) {
        R D = std::fabs(tf_ - tn);
        V3<R> tp = cb.eye + d * tn + G * R(0.5);
        auto sh = [&](const V3<R>& p, R* src) {
          if (!ph) return ebs.shade(p, src, &mg);
          // Phong branch (:523-546)
          R ka = 0, kd = 0, ks = 0, Io = 0, Is = 0;
          if (prm.apply_occlusion) { ka = R(prm.ka); Io = ebs.ambient_occlusion(p); }
          if (prm.apply_shadow) { kd = R(prm.kd); ks = R(P.ks); Is = ebs.shadows(p, &mg); }
          shade_lit_phong(P, G, p, Io, Is, ka, kd, ks, src);
        };
        march(tp, d, D, step, et, G, vol, tf.rgbt.data(), tf.size(), sh, o, &mg);
      }
      if (margin) { margin[(size_t(py) * W + px) * 2] = mg.a; margin[(size_t(py) * W + px) * 2 + 1] = mg.g; }
    }
}

}  // namespace orc
