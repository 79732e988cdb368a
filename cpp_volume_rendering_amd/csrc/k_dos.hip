// Directional occlusion shading with cone section tables (gfx950): the renderer of
// RC1PConeTracingDirOcclusionShading.
//
// Host side: the section tables of ConeGaussianSampler::ComputeConeIntegrationSteps
// (rc1pdosct/conegaussiansampler.cpp:212-414) built in FP64: sections are placed along
// the cone axis every d_sigma*sigma (per half interval); the Gaussian sigma doubles
// whenever the cone radius outgrows r_sigma*sigma, and before doubling the cone is
// split into 3 and then 7 narrower sub-cones (circle packings) as the packing limit
// allows.  Each section carries (interval to the next, mip level log2(sigma/sigma0),
// trapezoid half width, amplitude (pr*Ig)^2 / (pi R^2)) with pr the Gaussian's mass
// over the cone diameter by the midpoint rule.
//
// Device side: the march of rc1pdosct/ray_bbox_marching.comp:658-734; each opaque
// sample integrates the occlusion cone towards the eye and (optionally) the shadow
// cone towards the light through the tables (:124-333, :345-562): one Gaussian per
// section on the axis, then 3 then 7 sub-cones that inherit their parent's optical
// depth, trapezoid sums of amplitude * extinction, I = exp(-depth) averaged over the
// sub-cones (the 7-packing weights its ring by the adjacent weight).  The table row
// index is wave-uniform, so rows arrive through scalar loads and the level geometry
// stays in SGPRs.
#include "cvr_cone.hpp"
#include <cmath>
#include <cstring>
#include <vector>

namespace cvr {

// Per-cone constants of the device integration.
struct DosConeK {
  int n1, n3, n7, pad;
  float init, w, adj7, pad2;
  float ax[10][3];  // 3-packing axes, then 7-packing axes, (x, y, z) in the cone frame
};

struct DosK {
  DosConeK occ, sdw;
  const float4* occ_rows;
  const float4* sdw_rows;
  int apply_occ, apply_sdw, shadow_type;
  float spot_cos, ka, kd;
  float eye[3], lpos[3], lfwd[3], lup[3], lright[3];
};

// GetGaussianExtinction (ray_bbox_marching.comp:92-112) at a section's level: textureLod
// clamps the level to the chain, the border falloff uses the level as given.
template <int F>
__device__ __forceinline__ float dos_ext(const void* __restrict__ pyr, const PyrGeom& pg, f3 tp, float level) {
  const float lc = fminf(fmaxf(level, 0.f), float(pg.levels - 1));
  const int d1 = int(floorf(lc));
  const float fr = lc - float(d1);
  float a = pyr_tri<F>(pyr, pg, d1, tp);
  if (fr > 0.f) a = a * (1.f - fr) + pyr_tri<F>(pyr, pg, min(d1 + 1, pg.levels - 1), tp) * fr;
  return a * border_falloff(pg, tp, 0.5f * exp2f(-2.f * level));
}

// One cone through its section table; sub-cone axes are k*z + uy*y + vx*x.
template <int F>
__device__ __forceinline__ float dos_cone(const void* __restrict__ pyr, const PyrGeom& pg, const DosConeK& C,
                                          const float4* __restrict__ rows, f3 p, f3 k, f3 uy, f3 vx) {
  float od[7], lt[7];
  float track = C.init;
  od[0] = 0.f;
  lt[0] = 0.f;
  int i = 0;
  // the single-cone sections read one level each and their positions do not depend on
  // the reads: unrolled by 4, four reads are in flight at once and the sums keep their
  // order (256^3 @ 1024^2 with shadows: 37.7 -> 35.9 ms, the same frame bit for bit)
#pragma unroll 4
  for (; i < C.n1; ++i) {
    const float4 s = rows[i];
    const float at = dos_ext<F>(pyr, pg, p + k * track, s.y) * s.w;
    od[0] += (lt[0] + at) * s.z * C.w;
    lt[0] = at;
    track += s.x;
  }
  if (C.n3 + C.n7 == 0) return __expf(-od[0]);
  od[1] = od[2] = od[0];
  lt[1] = lt[2] = lt[0];
  f3 d3[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) d3[j] = k * C.ax[j][2] + uy * C.ax[j][1] + vx * C.ax[j][0];
  for (const int e = i + C.n3; i < e; ++i) {
    const float4 s = rows[i];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float at = dos_ext<F>(pyr, pg, p + d3[j] * track, s.y) * s.w;
      od[j] += (lt[j] + at) * s.z * C.w;
      lt[j] = at;
    }
    track += s.x;
  }
  if (C.n7 == 0) return (__expf(-od[0]) + __expf(-od[1]) + __expf(-od[2])) / 3.0f;
  // 3 -> 7: each outer cone inherits the depth of the 3-cone it came from, the centre
  // cone their mean (Cone7RayOcclusion, :127-137)
  const float ma = (od[2] + od[1] + od[0]) / 3.0f, mt = (lt[2] + lt[1] + lt[0]) / 3.0f;
  od[6] = od[5] = od[2];
  od[4] = od[3] = od[1];
  od[2] = od[1] = od[0];
  od[0] = ma;
  lt[6] = lt[5] = lt[2];
  lt[4] = lt[3] = lt[1];
  lt[2] = lt[1] = lt[0];
  lt[0] = mt;
  f3 d7[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) d7[j] = k * C.ax[3 + j][2] + uy * C.ax[3 + j][1] + vx * C.ax[3 + j][0];
  for (const int e = i + C.n7; i < e; ++i) {
    const float4 s = rows[i];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      const float at = dos_ext<F>(pyr, pg, p + d7[j] * track, s.y) * s.w;
      od[j] += (lt[j] + at) * s.z * C.w;
      lt[j] = at;
    }
    track += s.x;
  }
  const float ring = __expf(-od[1]) + __expf(-od[2]) + __expf(-od[3]) + __expf(-od[4]) + __expf(-od[5]) + __expf(-od[6]);
  return (__expf(-od[0]) + ring * C.adj7) / (1.0f + C.adj7 * 6.0f);
}

// Shadow cone of ShadowEvaluationKernel (:533-562).  `wp` is the world position.
template <int F>
__device__ __forceinline__ float dos_shadow(const void* __restrict__ pyr, const PyrGeom& pg, const DosK& K, f3 p,
                                            f3 wp) {
  if (K.shadow_type == 2) return dos_cone<F>(pyr, pg, K.sdw, K.sdw_rows, p, ld3(K.lfwd), ld3(K.lup), ld3(K.lright));
  const f3 k = norm3(ld3(K.lpos) - wp);
  if (K.shadow_type == 1 && dot3(k, ld3(K.lfwd)) < K.spot_cos) return 0.f;
  const f3 u1 = norm3(cross3(k, ld3(K.lright)));
  const f3 v1 = norm3(cross3(k, u1));
  return dos_cone<F>(pyr, pg, K.sdw, K.sdw_rows, p, k, v1, u1);
}

template <typename T, int F>
__global__ __launch_bounds__(256) void k_dos_render(RayFrame rf, VolGeom g, PyrGeom pg, const void* __restrict__ pyr,
                                                    DosK K, const float4* __restrict__ tf, int ntf, float step, float et,
                                                    PhongK ph, const float4* __restrict__ grad, float4* __restrict__ out,
                                                    unsigned long long* __restrict__ counters) {
  extern __shared__ float4 s_tf[];
  for (int i = threadIdx.x; i < ntf; i += blockDim.x) s_tf[i] = tf[i];
  __syncthreads();
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tiles_x = (rf.W + 15) >> 4;
  int tx, ty;
  super_tile(bid, tiles_x, (rf.H + 15) >> 4, 16, 16, tx, ty);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = tx * 16 + (w & 1) * 8 + (lane & 7);
  const int py = ty * 16 + (w >> 1) * 8 + (lane >> 3);
  if (px >= rf.W || py >= rf.H) return;
  const f3 dir = camera_ray(rf, px, py);
  const f3 eye = ld3(rf.eye);
  float tn, tf_;
  float4 dst = make_float4(0.f, 0.f, 0.f, 0.f);
  unsigned int nsamp = 0, nshade = 0;
  const float ka = K.apply_occ ? K.ka : 0.f, kd = K.apply_sdw ? K.kd : 0.f;
  const float inv_k = 1.0f / (ka + kd);
  if (ray_box(eye, dir, g.G, tn, tf_)) {
    const f3 right = norm3(cross3(dir, mk3(0.f, 1.f, 0.f)));
    const f3 up = norm3(cross3(mk3(-dir.x, -dir.y, -dir.z), right));
    const f3 half = mk3(0.5f * g.G[0], 0.5f * g.G[1], 0.5f * g.G[2]);
    const float D = fabsf(tf_ - tn);
    const f3 tp = eye + dir * tn + half;
    for (int j = 0; j < kMaxSamples; ++j) {
      const float s = float(j) * step;
      if (!(s < D)) break;
      const float h = fminf(step, D - s);
      const f3 p = tp + dir * (s + 0.5f * h);
      float4 src = sample_tf(s_tf, ntf, sample_volume_oct<T>(g, p));
      ++nsamp;
      if (src.w > 0.0f) {
        ++nshade;
        const f3 wp = p - half;
        float Io = 0.f, Is = 0.f;
        if (K.apply_occ) Io = dos_cone<F>(pyr, pg, K.occ, K.occ_rows, p, norm3(eye - wp), up, right);
        if (K.apply_sdw) Is = dos_shadow<F>(pyr, pg, K, p, wp);
        if (ph.on) {
          // Phong branch (:627-643): a sample whose gradient vanishes keeps its colour
          float dd, sp;
          if (phong_terms(grad, g, ph, p, eye, dd, sp)) {
            const float m = inv_k * (Io * ka + Is * kd * dd), s = Is * (K.apply_sdw ? ph.ks : 0.f) * sp;
            src.x = src.x * m + ph.spec[0] * s;
            src.y = src.y * m + ph.spec[1] * s;
            src.z = src.z * m + ph.spec[2] * s;
          }
        } else {
          const float m = inv_k * (Io * ka + Is * kd);
          src.x *= m; src.y *= m; src.z *= m;
        }
        const float a = 1.0f - __expf(-src.w * h);
        const float wgt = (1.0f - dst.w) * a;
        dst.x += wgt * src.x;
        dst.y += wgt * src.y;
        dst.z += wgt * src.z;
        dst.w += wgt;
        if (dst.w > et) break;
      }
    }
  }
  out[size_t(py) * rf.W + px] = dst;
  if (counters) {
    atomicAdd(&counters[0], (unsigned long long)nsamp);
    atomicAdd(&counters[1], (unsigned long long)nshade);
  }
}

// Object-space light cache (rc1pdosct/lightcachecomputation.comp:523-546): cell i
// holds (occlusion, shadow) at texture-space position (i + 1/2) * cell, 1 for a term
// that is off.  The occlusion cone turns towards the eye in the frame
// right = normalize(cross(-k, camera up)), up = normalize(cross(k, right)) (:279-291);
// the shadow cone is the renderer's own.  One lane per cell.
template <int F>
__global__ __launch_bounds__(256) void k_lc_dos(PyrGeom pg, const void* __restrict__ pyr, DosK K, f3 cam_up, int rx,
                                                int ry, int rz, f3 cell, f3 half, float2* __restrict__ lc) {
  const size_t n = size_t(rx) * ry * rz;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = int(i % size_t(rx)), y = int((i / size_t(rx)) % size_t(ry)), z = int(i / (size_t(rx) * ry));
  const f3 tp = mk3((float(x) + 0.5f) * cell.x, (float(y) + 0.5f) * cell.y, (float(z) + 0.5f) * cell.z);
  const f3 wp = tp - half;
  float Io = 1.f, Is = 1.f;
  if (K.apply_occ) {
    const f3 k = norm3(ld3(K.eye) - wp);
    const f3 vr = norm3(cross3(mk3(-k.x, -k.y, -k.z), cam_up));
    const f3 vu = norm3(cross3(k, vr));
    Io = dos_cone<F>(pyr, pg, K.occ, K.occ_rows, tp, k, vu, vr);
  }
  if (K.apply_sdw) Is = dos_shadow<F>(pyr, pg, K, tp, wp);
  lc[i] = make_float2(Io, Is);
}

// ------------------------------------------------------------------ host: section tables
namespace {
const double kPi = 3.14159265358979323846;
const double kDiv3 = 1.0 + 2.0 / std::sqrt(3.0);  // 3 circles packed in a circle: radius ratio
const double kDiv7 = 3.01;                         // 7 circles (the reference's rounded ratio)

struct Section {
  int ng;
  double pos, radius, sigma;
};

// Mass of N(0, sigma) over [-R, R], midpoint rule on segments of about 0.05.
double gauss_mass(double sigma, double R) {
  const int nt = int(std::ceil((2.0 * R) / 0.05));
  const double seg = (2.0 * R) / double(nt), x0 = -R + 0.5 * seg;
  const double norm = 1.0 / (std::sqrt(2.0 * kPi) * sigma);
  double S = 0.0;
  for (int i = 0; i < nt; ++i) {
    const double x = x0 + seg * double(i);
    S += norm * std::exp(-(x * x) / (2.0 * sigma * sigma)) * seg;
  }
  return S;
}

// Rotation of v by angle t about the unit axis a (Rodrigues), renormalised.
void rotate(const double v[3], double t, const double a[3], double out[3]) {
  const double c = std::cos(t), s = std::sin(t), d = a[0] * v[0] + a[1] * v[1] + a[2] * v[2];
  const double cr[3] = {a[1] * v[2] - a[2] * v[1], a[2] * v[0] - a[0] * v[2], a[0] * v[1] - a[1] * v[0]};
  double r[3], l = 0;
  for (int i = 0; i < 3; ++i) {
    r[i] = v[i] * c + cr[i] * s + a[i] * d * (1.0 - c);
    l += r[i] * r[i];
  }
  l = std::sqrt(l);
  for (int i = 0; i < 3; ++i) out[i] = r[i] / l;
}

struct Builder {
  double half_deg, r_sigma;
  int max_packing, ng = 1;
  std::vector<Section> secs;
  // Places a section at `pos` with sigma `sg`, splitting into more sub-cones (never
  // fewer than the previous section had) while the radius exceeds r_sigma * sg.
  bool add(double pos, double sg) {
    static const double div[3] = {1.0, kDiv3, kDiv7};
    static const int ngs[3] = {1, 3, 7};
    int k = ng == 1 ? 0 : (ng == 3 ? 1 : 2);
    for (;;) {
      ng = ngs[k];
      const double R = pos * std::tan((half_deg / div[k]) * kPi / 180.0);
      if (!(R > r_sigma * sg)) {
        secs.push_back(Section{ng, pos, R, sg});
        return true;
      }
      if (k >= max_packing) return false;
      ++k;
    }
  }
};
}  // namespace

// The clamps of ConeGaussianSampler's setters (conegaussiansampler.cpp:57-172).
static cvr_dos_cone clamp_cone(cvr_dos_cone c, double covered_default) {
  c.half_deg = std::min(std::max(c.half_deg, 0.5f), 89.5f);
  c.max_packing = std::min(std::max(c.max_packing, 0), 2);
  if (!(c.initial_step > 0)) c.initial_step = 0.f;
  const float cov = c.covered_dist > 0 ? c.covered_dist : float(covered_default);
  c.covered_dist = cov > 10.0f ? cov : 10.0f;
  c.d_sigma = std::min(std::max(c.d_sigma, 1.0f), 3.0f);
  c.r_sigma = std::min(std::max(c.r_sigma, 0.5f), 3.0f);
  return c;
}

// The schedule ends only if the Gaussians can outgrow the widest sub-cone allowed: the
// next section sits about d_sigma * sigma further, where that sub-cone's radius grows by
// d_sigma * sigma * tan(half / div); with d_sigma * tan(half / div) >= r_sigma, sigma
// would double without bound.
static bool dos_cone_converges(const cvr_dos_cone& C) {
  static const double div[3] = {1.0, kDiv3, kDiv7};
  const double t = std::tan((double(C.half_deg) / div[C.max_packing]) * kPi / 180.0);
  return double(C.d_sigma) * t < double(C.r_sigma);
}

int dos_build_table(const cvr_dos_cone& C, double sigma0, DosTable& T) {
  if (!(C.initial_step > 0) || !(C.initial_step < C.covered_dist)) {
    set_error("dos params: initial_step must be > 0 and below the covered distance");
    return CVR_ERR_INVALID;
  }
  if (!dos_cone_converges(C)) {
    set_error("dos params: d_sigma * tan(half angle / packing ratio) must stay below r_sigma "
              "(the section schedule would never close)");
    return CVR_ERR_INVALID;
  }
  Builder b{double(C.half_deg), double(C.r_sigma), C.max_packing};
  const double ds = double(C.d_sigma);
  double pos = double(C.initial_step), sg = sigma0;
  std::vector<double> iv;  // interval after each section
  while (!b.add(pos, sg)) sg *= 2.0;
  while (pos < double(C.covered_dist)) {
    double si = ds * sg;
    while (!b.add(pos + si + ds * sg, sg)) sg *= 2.0;
    si += ds * sg;
    iv.push_back(si);
    pos += si;
    if (iv.size() > (1u << 16)) {
      set_error("dos params: more than 65536 cone sections");
      return CVR_ERR_INVALID;
    }
  }
  iv.push_back(0.0);
  const size_t n = b.secs.size();
  T.rows.resize(n);
  T.n[0] = T.n[1] = T.n[2] = 0;
  for (size_t i = 0; i < n; ++i) {
    const Section& s = b.secs[i];
    T.n[s.ng == 1 ? 0 : (s.ng == 3 ? 1 : 2)] += 1;
    const double dint = i == 0 ? s.sigma * std::sqrt(2.0 * kPi) * 0.5 : iv[i - 1] * 0.5;
    const double pr = gauss_mass(s.sigma, s.radius);
    const double Ig = s.sigma * std::sqrt(2.0 * kPi);
    const double amp = (pr * pr) * (Ig * Ig) / (kPi * s.radius * s.radius);
    T.rows[i] = make_float4(float(iv[i]), float(std::log2(s.sigma / sigma0)), float(dint), float(amp));
  }
  // sub-cone axes (conegaussiansampler.cpp:219-252)
  const double z[3] = {0, 0, 1}, y[3] = {0, 1, 0};
  const double half = double(C.half_deg);
  double a[10][3];
  rotate(z, (half - half / kDiv3) * kPi / 180.0, y, a[0]);
  rotate(a[0], 30.0 * kPi / 180.0, z, a[0]);
  rotate(a[0], 120.0 * kPi / 180.0, z, a[1]);
  rotate(a[1], 120.0 * kPi / 180.0, z, a[2]);
  a[3][0] = 0; a[3][1] = 0; a[3][2] = 1;
  rotate(z, (half - half / kDiv7) * kPi / 180.0, y, a[4]);
  for (int j = 5; j < 10; ++j) rotate(a[j - 1], 60.0 * kPi / 180.0, z, a[j]);
  for (int j = 0; j < 10; ++j)
    for (int c = 0; c < 3; ++c) T.axes[j][c] = float(a[j][c]);
  T.adj7 = float(a[4][2]);
  T.init = C.initial_step;
  T.w = C.ui_weight;
  return CVR_OK;
}

static cvr_dos_cone clamp_cone(cvr_dos_cone c, double covered_default);

int check_dos_params(const cvr_dos_params& P) {
  if (!dos_cone_converges(clamp_cone(P.occ, 100.0)) || !dos_cone_converges(clamp_cone(P.sdw, 100.0))) {
    set_error("dos params: d_sigma * tan(half angle / packing ratio) must stay below r_sigma "
              "(the section schedule would never close)");
    return CVR_ERR_INVALID;
  }
  if (P.apply_occlusion == 0 && P.apply_shadow == 0) {
    set_error("dos params: occlusion or shadow must be enabled (the shading divides by ka+kd)");
    return CVR_ERR_INVALID;
  }
  if (!(P.sigma0 > 0) || P.n_levels > kMaxLevels || P.shadow_type < 0 || P.shadow_type > 2 || !(P.et_threshold > 0) ||
      !(P.occ.initial_step > 0) || !(P.sdw.initial_step > 0) || P.density_precision < 0 || P.density_precision > 2) {
    set_error("dos params: sigma0 > 0, n_levels <= 12, shadow_type in {0,1,2}, et_threshold > 0, initial steps > 0, "
              "density_precision 0, 1 or 2");
    return CVR_ERR_INVALID;
  }
  return CVR_OK;
}

int prepare_dos(cvr_ctx* c, const cvr_dos_params& P, hipStream_t st) {
  int rc = check_dos_params(P);
  if (rc) return rc;
  const int bx = P.base_res[0] > 0 ? P.base_res[0] : c->nx;
  const int by = P.base_res[1] > 0 ? P.base_res[1] : c->ny;
  const int bz = P.base_res[2] > 0 ? P.base_res[2] : c->nz;
  rc = build_pyramid(c, bx, by, bz, P.sigma0, P.n_levels, P.density_precision, st);
  if (rc) return rc;
  // SetCoveredDistance(diagonal * 0.50f / 0.75f) in float (dosrcrenderer.cpp:112-113)
  const float diag = volume_diagonal(c);
  const cvr_dos_cone oc = clamp_cone(P.occ, diag * 0.50f), sc = clamp_cone(P.sdw, diag * 0.75f);
  if ((rc = dos_build_table(oc, P.sigma0, c->dos_occ))) return rc;
  if ((rc = dos_build_table(sc, P.sigma0, c->dos_sdw))) return rc;
  const size_t n = c->dos_occ.rows.size() + c->dos_sdw.rows.size();
  if (n > c->dos_rows_cap) {
    if (c->d_dos_rows) CVR_HIP_CHECK(hipFree(c->d_dos_rows));
    c->d_dos_rows = nullptr;
    c->dos_rows_cap = 0;
    if (hipMalloc(&c->d_dos_rows, n * sizeof(float4)) != hipSuccess) {
      set_error("dos: out of device memory");
      return CVR_ERR_NOMEM;
    }
    c->dos_rows_cap = n;
  }
  std::vector<float4> all(c->dos_occ.rows);
  all.insert(all.end(), c->dos_sdw.rows.begin(), c->dos_sdw.rows.end());
  CVR_HIP_CHECK(hipMemcpy(c->d_dos_rows, all.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  return CVR_OK;
}

static DosConeK cone_k(const DosTable& T) {
  DosConeK k{};
  k.n1 = T.n[0];
  k.n3 = T.n[1];
  k.n7 = T.n[2];
  k.init = T.init;
  k.w = T.w;
  k.adj7 = T.adj7;
  std::memcpy(k.ax, T.axes, sizeof(k.ax));
  return k;
}

int launch_dos(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st) {
  const cvr_dos_params& P = c->dos;
  int rc = prepare_dos(c, P, st);
  if (rc) return rc;
  DosK K{};
  K.occ = cone_k(c->dos_occ);
  K.sdw = cone_k(c->dos_sdw);
  K.occ_rows = c->d_dos_rows;
  K.sdw_rows = c->d_dos_rows + c->dos_occ.rows.size();
  K.apply_occ = P.apply_occlusion;
  K.apply_sdw = P.apply_shadow;
  K.shadow_type = P.shadow_type;
  K.spot_cos = float(std::cos(double(P.spot_max_deg) * kPi / 180.0));
  K.ka = P.ka;
  K.kd = P.kd;
  for (int i = 0; i < 3; ++i) {
    K.eye[i] = rf.eye[i];
    // without a light: the camera's own light (renderingmanager.cpp:421-426)
    K.lpos[i] = light ? light->position[i] : rf.eye[i];
    K.lfwd[i] = light ? light->forward[i] : -rf.f[i];
    K.lup[i] = light ? light->up[i] : rf.u[i];
    K.lright[i] = light ? light->right[i] : rf.s[i];
  }
  VolGeom g = geom_of(c);
  const float step = P.step_size > 0 ? P.step_size : default_step(c);
  const int tiles = ((rf.W + 15) / 16) * ((rf.H + 15) / 16);
  const size_t shm = size_t(c->tf_size) * sizeof(float4);
  if (shm > 64 * 1024) { set_error("dos renderer: transfer function larger than 4096 entries"); return CVR_ERR_UNSUPPORTED; }
  PhongK ph;
  if ((rc = phong_setup(c, P.apply_phong, P.gradient_type, P.ks, P.shininess, P.specular, rf, light, ph, st))) return rc;
  if (P.light_cache) {
    // PreComputeLightCache (dosrcrenderer.cpp:555-651), then the object-space march
    if ((rc = lc_prepare(c, P.light_cache_res))) return rc;
    const int rx = c->lc_res[0], ry = c->lc_res[1], rz = c->lc_res[2];
    const int dims[3] = {c->nx, c->ny, c->nz}, res[3] = {rx, ry, rz};
    const float sc[3] = {c->sx, c->sy, c->sz};
    float cs[3];
    for (int i = 0; i < 3; ++i) cs[i] = sc[i] * (float(dims[i]) / float(res[i]));  // VolumeScales * (dims / res)
    const size_t n = size_t(rx) * ry * rz;
    const f3 cam_up{rf.u[0], rf.u[1], rf.u[2]};
    const f3 half{0.5f * g.G[0], 0.5f * g.G[1], 0.5f * g.G[2]};
    const f3 cell{cs[0], cs[1], cs[2]};
    with_fmt(c->pyr.fmt, [&](auto f) {
      hipLaunchKernelGGL((k_lc_dos<decltype(f)::value>), dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, c->pyr,
                         c->d_pyr_oct, K, cam_up, rx, ry, rz, cell, half, c->d_lc);
    });
    CVR_HIP_CHECK(hipGetLastError());
    return launch_lc_render(c, rf, P.apply_occlusion, P.apply_shadow, P.ka, P.kd, step, P.et_threshold, ph, out, st);
  }
  unsigned long long* cnt = c->collect_stats ? c->d_counters : nullptr;
  float4* o4 = reinterpret_cast<float4*>(out);
  with_types(c->dtype, c->pyr.fmt, [&](auto t, auto f) {
    using T = typename decltype(t)::type;
    hipLaunchKernelGGL((k_dos_render<T, decltype(f)::value>), dim3(tiles), dim3(256), shm, st, rf, g, c->pyr,
                       c->d_pyr_oct, K, c->d_tf_rgbt, c->tf_size, step, P.et_threshold, ph, c->d_grad, o4, cnt);
  });
  CVR_HIP_CHECK(hipGetLastError());
  return CVR_OK;
}

}  // namespace cvr
