// CPU ORACLE — cone-traced lighting (test infrastructure only).
//
// Gaussian extinction mip pyramid (rc1pdosct/extcoefvolumegenerator.cpp:230-408 and
// glslextgen/{gen_extcoefvol_anysize,gen_extcoefvol_anysize_mmlevel,backtotau}.comp),
// its LINEAR_MIPMAP_LINEAR lookups with the border falloff of
// rc1pdosct/ray_bbox_marching.comp:92-112, the cone-traced directional occlusion +
// shadow renderer built on it (the configuration with 8 occlusion cones and one
// shadow cone, each marched over a fixed number of steps), the surrogate's input
// features, and an FP64 multilayer perceptron (forward and backward) used to check
// the GPU training kernels.
#pragma once
#include "oracle.hpp"

namespace orc {

static const double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- pyramid
struct Pyramid {
  int levels = 0;
  std::vector<int> nx, ny, nz;
  std::vector<std::vector<double>> tau;  // extinction per level (after backtotau)
  std::vector<std::vector<double>> op;   // opacity per level (before backtotau)
  double G[3] = {0, 0, 0};
};

// GL mip chain sizes: max(1, n >> i) per axis, down to 1^3.
inline int mip_count(int bx, int by, int bz) {
  int m = std::max(bx, std::max(by, bz)), c = 1;
  while (m > 1) { m >>= 1; ++c; }
  return c;
}

// gen_extcoefvol_anysize.comp:16-77 — level 0 at base resolution (bx,by,bz), sigma S0,
// opacity from the TF RGBA table of the trilinear volume sample, 0 outside [0,1]^3.
inline void pyramid_level0(const Volume& vol, const TF& tf, int bx, int by, int bz, double S0, const double G[3],
                           std::vector<double>& out) {
  const double vs[3] = {G[0] / bx, G[1] / by, G[2] / bz};
  out.assign(size_t(bx) * by * bz, 0.0);
  const int T = tf.size();
#pragma omp parallel for schedule(dynamic, 1)
  for (int z = 0; z < bz; ++z)
    for (int y = 0; y < by; ++y)
      for (int x = 0; x < bx; ++x) {
        double gp[3] = {(x + 0.5) * vs[0], (y + 0.5) * vs[1], (z + 0.5) * vs[2]};
        double swc = 0, sw = 0;
        for (int i = -3; i <= 3; ++i)
          for (int j = -3; j <= 3; ++j)
            for (int k = -3; k <= 3; ++k) {
              double f[3] = {i * S0, j * S0, k * S0};
              double wk = (S0 * S0 * S0) * std::exp(-(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) / (2.0 * S0 * S0));
              double t[3] = {(gp[0] + f[0]) / G[0], (gp[1] + f[1]) / G[1], (gp[2] + f[2]) / G[2]};
              double ck = 0;
              if (!(t[0] < 0 || t[1] < 0 || t[2] < 0 || t[0] > 1 || t[1] > 1 || t[2] > 1)) {
                double d = tex3d_volume<double>(vol, V3<double>(t[0], t[1], t[2]));
                double c[4];
                tex1d_tf<double>(tf.rgba.data(), T, d, c);
                ck = c[3];
              }
              swc += wk * ck;
              sw += wk;
            }
        out[size_t(x) + size_t(bx) * (y + size_t(by) * z)] = swc / sw;
      }
}

// gen_extcoefvol_anysize_mmlevel.comp:15-78 — level i from the trilinear lookups of level i-1.
inline void pyramid_next(const std::vector<double>& prev, int px, int py, int pz, int nx, int ny, int nz, double Si,
                         const double G[3], std::vector<double>& out) {
  const double lvs[3] = {G[0] / nx, G[1] / ny, G[2] / nz};
  out.assign(size_t(nx) * ny * nz, 0.0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        double gp[3] = {(x + 0.5) * lvs[0], (y + 0.5) * lvs[1], (z + 0.5) * lvs[2]};
        double swc = 0, sw = 0;
        for (int i = -3; i <= 3; ++i)
          for (int j = -3; j <= 3; ++j)
            for (int k = -3; k <= 3; ++k) {
              double f[3] = {i * Si, j * Si, k * Si};
              double wk = (Si * Si * Si) * std::exp(-(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) / (2.0 * Si * Si));
              double t[3] = {(gp[0] + f[0]) / G[0], (gp[1] + f[1]) / G[1], (gp[2] + f[2]) / G[2]};
              double ck = 0;
              if (!(t[0] < 0 || t[1] < 0 || t[2] < 0 || t[0] > 1 || t[1] > 1 || t[2] > 1))
                ck = tex3d_grid<double, double>(prev.data(), px, py, pz, V3<double>(t[0], t[1], t[2]));
              swc += wk * ck;
              sw += wk;
            }
        out[size_t(x) + size_t(nx) * (y + size_t(ny) * z)] = swc / sw;
      }
}

// extcoefvolumegenerator.cpp:230-367 + backtotau.comp:11-35.  `nlev` <= 0 builds the
// whole chain down to 1^3 as the reference does.
inline void pyramid_build(const Volume& vol, const TF& tf, int bx, int by, int bz, double S0, int nlev, Pyramid& P) {
  P.G[0] = vol.nx * vol.sx; P.G[1] = vol.ny * vol.sy; P.G[2] = vol.nz * vol.sz;
  int full = mip_count(bx, by, bz);
  P.levels = nlev > 0 ? std::min(nlev, full) : full;
  P.nx.clear(); P.ny.clear(); P.nz.clear(); P.op.clear(); P.tau.clear();
  for (int i = 0; i < P.levels; ++i) {
    P.nx.push_back(std::max(1, bx >> i));
    P.ny.push_back(std::max(1, by >> i));
    P.nz.push_back(std::max(1, bz >> i));
  }
  P.op.resize(P.levels);
  pyramid_level0(vol, tf, bx, by, bz, S0, P.G, P.op[0]);
  for (int i = 1; i < P.levels; ++i)
    pyramid_next(P.op[i - 1], P.nx[i - 1], P.ny[i - 1], P.nz[i - 1], P.nx[i], P.ny[i], P.nz[i],
                 S0 * std::pow(2.0, double(i)), P.G, P.op[i]);
  P.tau.resize(P.levels);
  for (int i = 0; i < P.levels; ++i) {
    P.tau[i].resize(P.op[i].size());
    for (size_t j = 0; j < P.op[i].size(); ++j) P.tau[i][j] = -1.0 * std::log(1.0 - P.op[i][j]);
  }
}

// textureLod with LINEAR_MIPMAP_LINEAR (OpenGL 4.3, 8.14.2.1): lod clamped to the
// chain, lerp of the trilinear samples of floor(lod) and floor(lod)+1.
template <class R> inline R pyr_lod(const Pyramid& P, const V3<R>& uvw, R lod) {
  R q = R(P.levels - 1);
  if (lod < 0) lod = 0;
  if (lod > q) lod = q;
  int d1 = int(std::floor(lod));
  R fr = lod - R(d1);
  R a = tex3d_grid<R, double>(P.tau[d1].data(), P.nx[d1], P.ny[d1], P.nz[d1], uvw);
  if (fr == 0 || d1 + 1 >= P.levels) return a;
  int d2 = d1 + 1;
  R b = tex3d_grid<R, double>(P.tau[d2].data(), P.nx[d2], P.ny[d2], P.nz[d2], uvw);
  return a * (1 - fr) + b * fr;
}

// GetGaussianExtinction, ray_bbox_marching.comp:92-112 (CONSIDER_BORDERS defined).
template <class R> inline R gauss_ext(const Pyramid& P, const V3<R>& tp, R lod) {
  V3<R> G(R(P.G[0]), R(P.G[1]), R(P.G[2]));
  R rg = pyr_lod<R>(P, tp / G, lod);
  if (tp.x < 0 || tp.x > G.x || tp.y < 0 || tp.y > G.y || tp.z < 0 || tp.z > G.z) {
    R sg = std::pow(R(2), lod);
    V3<R> c = vclamp(tp, V3<R>(0, 0, 0), G) - tp;
    R dist = c.x * c.x + c.y * c.y + c.z * c.z;
    rg = rg * std::exp(-dist / (R(2) * sg * sg));
  }
  return rg;
}

// ---------------------------------------------------------------- cone tracer
struct ConeParams {
  int n_steps = 64;
  int n_occ = 8;              // 1 axial + (n_occ-1) on a ring
  double occ_half_deg = 30;   // half angle of each occlusion cone
  double ring_deg = 60;       // angle of the ring cones' axes to the occlusion axis
  double sdw_half_deg = 10;   // half angle of the shadow cone
  double t0 = 1.0;            // first distance (world units)
  double occ_max = 0, sdw_max = 0;  // march lengths (world units)
  double r_sigma = 2.0;       // cone radius / Gaussian sigma at the chosen level
  double sigma0 = 1.0;        // sigma of level 0 (world units)
  int apply_occ = 1, apply_sdw = 1;
  double ka = 0.5, kd = 0.5;
  double light_fwd[3] = {0, 0, 1};
};

// One cone's step schedule: geometric distances t_j = t0 q^j, q = (tmax/t0)^(1/n); each
// step samples its midpoint at the level whose sigma matches the cone radius there.
struct ConeStep { double tm, dt, lod; };
inline std::vector<ConeStep> cone_schedule(int n, double t0, double tmax, double half_deg, double r_sigma,
                                           double sigma0, int levels) {
  std::vector<ConeStep> s(n);
  double q = std::pow(tmax / t0, 1.0 / n), th = std::tan(half_deg * kPi / 180.0);
  for (int j = 0; j < n; ++j) {
    double a = t0 * std::pow(q, double(j)), b = t0 * std::pow(q, double(j + 1));
    double tm = 0.5 * (a + b), r = tm * th;
    double lod = std::log2(r / (r_sigma * sigma0));
    lod = std::min(std::max(lod, 0.0), double(levels - 1));
    s[j] = ConeStep{tm, b - a, lod};
  }
  return s;
}

template <class R> inline R cone_transmittance(const Pyramid& P, const std::vector<ConeStep>& sch, const V3<R>& p,
                                               const V3<R>& axis) {
  R od = 0;
  for (const ConeStep& s : sch) od += gauss_ext<R>(P, p + axis * R(s.tm), R(s.lod)) * R(s.dt);
  return std::exp(-od);
}

// Occlusion frame of a camera ray (gt_ray_marching.comp:396-399): k = -dir,
// right = normalize(cross(dir, (0,1,0))), up = normalize(cross(-dir, right)).
template <class R> inline void occ_frame(const V3<R>& dir, V3<R>& k, V3<R>& up, V3<R>& right) {
  right = normalize(cross(dir, V3<R>(0, 1, 0)));
  up = normalize(cross(-dir, right));
  k = -dir;
}

template <class R> struct ConeTracer {
  const Pyramid& P;
  ConeParams p;
  std::vector<ConeStep> so, ss;
  ConeTracer(const Pyramid& P_, const ConeParams& p_) : P(P_), p(p_) {
    so = cone_schedule(p.n_steps, p.t0, p.occ_max, p.occ_half_deg, p.r_sigma, p.sigma0, P.levels);
    ss = cone_schedule(p.n_steps, p.t0, p.sdw_max, p.sdw_half_deg, p.r_sigma, p.sigma0, P.levels);
  }
  R occlusion(const V3<R>& pos, const V3<R>& k, const V3<R>& up, const V3<R>& right) const {
    R num = cone_transmittance<R>(P, so, pos, k), den = 1;
    R ce = R(std::cos(p.ring_deg * kPi / 180.0)), se = R(std::sin(p.ring_deg * kPi / 180.0));
    for (int j = 0; j + 1 < p.n_occ; ++j) {
      double phi = 2.0 * kPi * j / (p.n_occ - 1);
      V3<R> ax = k * ce + (up * R(std::cos(phi)) + right * R(std::sin(phi))) * se;
      num += ce * cone_transmittance<R>(P, so, pos, ax);
      den += ce;
    }
    return num / den;
  }
  R shadow(const V3<R>& pos) const {
    V3<R> L = normalize(V3<R>(R(p.light_fwd[0]), R(p.light_fwd[1]), R(p.light_fwd[2])));
    return cone_transmittance<R>(P, ss, pos, L);
  }
};

template <class R>
inline void render_cone(const Volume& vol, const TF& tf, const Pyramid& P, const ConeParams& cp, const CamIn& cam,
                        int W, int H, R step, R et, R* out, double* margin, int stride = 1) {
  ConeTracer<R> ct(P, cp);
  CamBasis<R> cb = cam_basis<R>(cam, W, H);
  V3<R> G(R(vol.nx * vol.sx), R(vol.ny * vol.sy), R(vol.nz * vol.sz));
  R ka = cp.apply_occ ? R(cp.ka) : R(0), kd = cp.apply_sdw ? R(cp.kd) : R(0);
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
        V3<R> k, up, right;
        occ_frame(d, k, up, right);
        auto sh = [&](const V3<R>& p, R* src) {
          R Io = cp.apply_occ ? ct.occlusion(p, k, up, right) : R(0);
          R Is = cp.apply_sdw ? ct.shadow(p) : R(0);
          for (int c = 0; c < 3; ++c) src[c] = (R(1) / (ka + kd)) * (src[c] * Io * ka + src[c] * Is * kd);
        };
        march(tp, d, D, step, et, G, vol, tf.rgbt.data(), tf.size(), sh, o, &mg);
      }
      if (margin) { margin[(size_t(py) * W + px) * 2] = mg.a; margin[(size_t(py) * W + px) * 2 + 1] = mg.g; }
    }
}

// ---------------------------------------------------------------- surrogate features
// 16 inputs of the learned occlusion/shadow surrogate at sample position p with
// occlusion axis k and light direction L.  Occlusion: 4 geometric segments over
// [t0, occ_max]; for segment i (length b-a, mid distance D = sqrt(ab)) the opacity
// 1 - exp(-tau * (b-a)) of the pyramid read on the axis at p + k D at the level of
// the axial cone's radius D tan(half) (inputs 2i), and read at p + k D cos(ring) at
// the level of the ring's radius D sin(ring) (inputs 2i+1).  Shadow: 7 segments over
// [t0, sdw_max] along L at the level of the shadow cone's radius (inputs 8..14).
// Input 15: the opacity of level 0 at p over one unit.  Levels are rounded to whole
// levels (one trilinear read per input).
constexpr int kNFeat = 16, kNOcc = 4, kNSdw = 7;
template <class R>
inline void nn_features(const Pyramid& P, const ConeParams& cp, const V3<R>& p, const V3<R>& k, const V3<R>& L,
                        R* f) {
  const double d2r = kPi / 180.0;
  const double tha = std::tan(cp.occ_half_deg * d2r), cr = std::cos(cp.ring_deg * d2r), sr = std::sin(cp.ring_deg * d2r);
  const double ts = std::tan(cp.sdw_half_deg * d2r), rs = cp.r_sigma * cp.sigma0, top = double(P.levels - 1);
  auto lvl = [&](double r) { return std::min(std::max(std::round(std::log2(r / rs)), 0.0), top); };
  for (int i = 0; i < kNOcc; ++i) {
    double a = cp.t0 * std::pow(cp.occ_max / cp.t0, double(i) / kNOcc);
    double b = cp.t0 * std::pow(cp.occ_max / cp.t0, double(i + 1) / kNOcc);
    double D = std::sqrt(a * b);
    R t = gauss_ext<R>(P, p + k * R(D), R(lvl(D * tha)));
    f[2 * i] = R(1) - std::exp(-t * R(b - a));
    t = gauss_ext<R>(P, p + k * R(D * cr), R(lvl(D * sr)));
    f[2 * i + 1] = R(1) - std::exp(-t * R(b - a));
  }
  for (int i = 0; i < kNSdw; ++i) {
    double a = cp.t0 * std::pow(cp.sdw_max / cp.t0, double(i) / kNSdw);
    double b = cp.t0 * std::pow(cp.sdw_max / cp.t0, double(i + 1) / kNSdw);
    double D = std::sqrt(a * b);
    R t = gauss_ext<R>(P, p + L * R(D), R(lvl(D * ts)));
    f[8 + i] = R(1) - std::exp(-t * R(b - a));
  }
  f[15] = R(1) - std::exp(-gauss_ext<R>(P, p, R(0)));
}

// Network-free estimate from the same inputs: transmittance products along the axial
// cone and along the ring, weighted like the exact tracer's cones, and along the light.
// The network predicts the residual to it.
inline void nn_baseline(const ConeParams& cp, const double* f, double& occ, double& sdw) {
  double ta = 1, tr = 1, ts = 1;
  for (int i = 0; i < kNOcc; ++i) { ta *= 1 - f[2 * i]; tr *= 1 - f[2 * i + 1]; }
  for (int i = 0; i < kNSdw; ++i) ts *= 1 - f[2 * kNOcc + i];
  const double wr = (cp.n_occ - 1) * std::cos(cp.ring_deg * kPi / 180.0);
  occ = (ta + wr * tr) / (1 + wr);
  sdw = ts;
}

// ---------------------------------------------------------------- MLP (FP64)
// Layers 16 -> 64 -> 64 -> 64 -> 16 (2 outputs used), ReLU on hidden layers, linear
// output.  Parameters flattened as W1[64][16], b1[64], W2[64][64], b2, W3[64][64], b3,
// W4[16][64], b4[16] (row = output feature).
constexpr int kH = 64, kOut = 16;
constexpr int kNParams = kH * kNFeat + kH + kH * kH + kH + kH * kH + kH + kOut * kH + kOut;
struct MlpLayout {
  int off_w[4], off_b[4], in[4], out[4];
  MlpLayout() {
    const int ins[4] = {kNFeat, kH, kH, kH}, outs[4] = {kH, kH, kH, kOut};
    int o = 0;
    for (int l = 0; l < 4; ++l) {
      in[l] = ins[l]; out[l] = outs[l];
      off_w[l] = o; o += outs[l] * ins[l];
      off_b[l] = o; o += outs[l];
    }
  }
};

// Forward for one sample; `acts` (optional) receives the 4 post-activation vectors.
inline void mlp_forward(const double* prm, const double* x, double* y, std::vector<double>* acts = nullptr) {
  MlpLayout L;
  std::vector<double> a(x, x + kNFeat), b;
  if (acts) acts->assign(a.begin(), a.end());
  for (int l = 0; l < 4; ++l) {
    b.assign(L.out[l], 0.0);
    for (int o = 0; o < L.out[l]; ++o) {
      double s = prm[L.off_b[l] + o];
      for (int i = 0; i < L.in[l]; ++i) s += prm[L.off_w[l] + o * L.in[l] + i] * a[i];
      b[o] = (l < 3) ? std::max(s, 0.0) : s;
    }
    a.swap(b);
    if (acts) acts->insert(acts->end(), a.begin(), a.end());
  }
  for (int o = 0; o < kOut; ++o) y[o] = a[o];
}

// Mean-squared-error loss over the first 2 outputs, L = sum_n sum_c (y - t)^2 / (2 N),
// and its gradient with respect to every parameter.
inline double mlp_loss_grad(const double* prm, const double* X, const double* T, int n, double* grad) {
  MlpLayout L;
  std::fill(grad, grad + kNParams, 0.0);
  double loss = 0;
  for (int s = 0; s < n; ++s) {
    std::vector<double> acts;
    double y[kOut];
    mlp_forward(prm, X + size_t(s) * kNFeat, y, &acts);
    // acts layout: x(16) h1(64) h2(64) h3(64) y(16)
    int aoff[5] = {0, kNFeat, kNFeat + kH, kNFeat + 2 * kH, kNFeat + 3 * kH};
    std::vector<double> g(kOut, 0.0), gp;
    for (int c = 0; c < 2; ++c) {
      double e = y[c] - T[size_t(s) * 2 + c];
      loss += e * e / (2.0 * n);
      g[c] = e / n;
    }
    for (int l = 3; l >= 0; --l) {
      const double* ain = acts.data() + aoff[l];
      for (int o = 0; o < L.out[l]; ++o) {
        grad[L.off_b[l] + o] += g[o];
        for (int i = 0; i < L.in[l]; ++i) grad[L.off_w[l] + o * L.in[l] + i] += g[o] * ain[i];
      }
      if (l == 0) break;
      gp.assign(L.in[l], 0.0);
      for (int i = 0; i < L.in[l]; ++i) {
        double s2 = 0;
        for (int o = 0; o < L.out[l]; ++o) s2 += prm[L.off_w[l] + o * L.in[l] + i] * g[o];
        gp[i] = ain[i] > 0 ? s2 : 0.0;  // ReLU of the layer below
      }
      g.swap(gp);
    }
  }
  return loss;
}

// bf16 round to nearest even of an fp32 value (the GPU rounds fp32 sums and inputs).
inline double bf16_round(double x) {
  float f = float(x);
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
  std::memcpy(&f, &u, 4);
  return double(f);
}

// Forward with the rounding points of a bf16 network with wide accumulation: the
// inputs, the weights and every hidden activation rounded to bf16 (after the ReLU; the
// order does not matter, bf16(max(x, 0)) = max(bf16(x), 0)), biases as given, products
// summed in FP64.  The output layer stays unrounded.
inline void mlp_forward_bf16(const double* prm, const double* x, double* y) {
  MlpLayout L;
  std::vector<double> a(kNFeat), b;
  for (int i = 0; i < kNFeat; ++i) a[i] = bf16_round(x[i]);
  for (int l = 0; l < 4; ++l) {
    b.assign(L.out[l], 0.0);
    for (int o = 0; o < L.out[l]; ++o) {
      double s = prm[L.off_b[l] + o];
      for (int i = 0; i < L.in[l]; ++i) s += bf16_round(prm[L.off_w[l] + o * L.in[l] + i]) * a[i];
      b[o] = (l < 3) ? bf16_round(std::max(s, 0.0)) : s;
    }
    a.swap(b);
  }
  for (int o = 0; o < kOut; ++o) y[o] = a[o];
}

// ---------------------------------------------------------------- surrogate-shaded frame
// The learned renderer's frame: the plain march of render_cone, where each opaque
// sample's (occlusion, shadow) is the network's residual added to the network-free
// estimate of its 16 inputs (nn_features, nn_baseline), clamped to [0, 1], and shades
// the sample as rc1pdosct/ray_bbox_marching.comp:607-656 does with the exact cone
// integrals: rgb <- rgb (Io ka + Is kd) / (ka + kd) (ka, kd zero for a term that is
// off).  A sample whose inputs leave [lo, hi] (the training range widened by the
// margin) takes the exact tracer's answers instead (ConeTracer).  `bf16` selects the
// network's rounding points (mlp_forward_bf16) or plain FP64 (mlp_forward).  counts[0]
// receives the shaded samples, counts[1] the samples the exact tracer shaded.
struct NnNet {
  const double* prm;  // kNParams, flattened as mlp_forward reads them
  const double* lo;   // kNFeat
  const double* hi;   // kNFeat
  int bf16;
};

template <class R>
inline void render_nn(const Volume& vol, const TF& tf, const Pyramid& P, const ConeParams& cp, const NnNet& net,
                      const CamIn& cam, int W, int H, R step, R et, R* out, double* margin, int stride,
                      long long* counts) {
  ConeTracer<R> ct(P, cp);
  CamBasis<R> cb = cam_basis<R>(cam, W, H);
  V3<R> G(R(vol.nx * vol.sx), R(vol.ny * vol.sy), R(vol.nz * vol.sz));
  V3<R> L = normalize(V3<R>(R(cp.light_fwd[0]), R(cp.light_fwd[1]), R(cp.light_fwd[2])));
  const double ka = cp.apply_occ ? cp.ka : 0.0, kd = cp.apply_sdw ? cp.kd : 0.0;
  long long n_shaded = 0, n_exact = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : n_shaded, n_exact)
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
        V3<R> k, up, right;
        occ_frame(d, k, up, right);
        auto sh = [&](const V3<R>& p, R* src) {
          R fr[kNFeat];
          nn_features<R>(P, cp, p, k, L, fr);
          double f[kNFeat], y[kOut], bo, bs;
          bool ood = false;
          for (int c = 0; c < kNFeat; ++c) {
            f[c] = double(fr[c]);
            ood |= (f[c] < net.lo[c]) | (f[c] > net.hi[c]);
          }
          double Io, Is;
          ++n_shaded;
          if (ood) {
            ++n_exact;
            Io = cp.apply_occ ? double(ct.occlusion(p, k, up, right)) : 0.0;
            Is = cp.apply_sdw ? double(ct.shadow(p)) : 0.0;
          } else {
            if (net.bf16) mlp_forward_bf16(net.prm, f, y);
            else mlp_forward(net.prm, f, y);
            nn_baseline(cp, f, bo, bs);
            Io = y[0] + bo;
            Is = y[1] + bs;
          }
          Io = std::min(std::max(Io, 0.0), 1.0);
          Is = std::min(std::max(Is, 0.0), 1.0);
          const R m = R((Io * ka + Is * kd) / (ka + kd));
          for (int c = 0; c < 3; ++c) src[c] = src[c] * m;
        };
        march(tp, d, D, step, et, G, vol, tf.rgbt.data(), tf.size(), sh, o, &mg);
      }
      if (margin) { margin[(size_t(py) * W + px) * 2] = mg.a; margin[(size_t(py) * W + px) * 2 + 1] = mg.g; }
    }
  if (counts) { counts[0] = n_shaded; counts[1] = n_exact; }
}

}  // namespace orc
