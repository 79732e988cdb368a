// Cone-traced directional occlusion + shadows over the Gaussian extinction pyramid
// (the exact tracer), gfx950.
//
// The march, transfer function and compositing are those of
// rc1pdosct/ray_bbox_marching.comp:658-734; each opaque sample is shaded with
// rgb * (I_occ*ka + I_sdw*kd) / (ka + kd) (:607-656).  I_occ averages 1 axial + 7
// ring cones around the direction back to the eye (cosine weights, frame of
// gt_ray_marching.comp:396-399), I_sdw is one cone towards the light; each cone is
// marched over a fixed number of geometric steps and every step reads the pyramid
// level whose Gaussian sigma matches the cone radius there (LINEAR_MIPMAP_LINEAR +
// the border falloff of GetGaussianExtinction, :92-112).
//
// Per step the level index is wave-uniform, so the level geometry stays in SGPRs;
// the step schedule is read with scalar loads from a small device table.
#include "cvr_cone.hpp"
#include <cmath>
#include <cstring>
#include <vector>

namespace cvr {

template <typename T, int F>
__global__ __launch_bounds__(256) void k_cone_render(RayFrame rf, VolGeom g, PyrGeom pg, const void* __restrict__ pyr,
                                                     ConeK K, const float4* __restrict__ tf, int ntf, float step,
                                                     float et, float4* __restrict__ out,
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
    f3 k, up, right;
    occ_frame(dir, k, up, right);
    const f3 L = ld3(K.L);
    const float D = fabsf(tf_ - tn);
    const f3 tp = eye + dir * tn + mk3(0.5f * g.G[0], 0.5f * g.G[1], 0.5f * g.G[2]);
    for (int j = 0; j < kMaxSamples; ++j) {
      const float s = float(j) * step;
      if (!(s < D)) break;
      const float h = fminf(step, D - s);
      const f3 p = tp + dir * (s + 0.5f * h);
      float4 src = sample_tf(s_tf, ntf, sample_volume_oct<T>(g, p));
      ++nsamp;
      if (src.w > 0.0f) {
        ++nshade;
        const float Io = K.apply_occ ? ct_occlusion<F>(pyr, pg, K, p, k, up, right) : 0.f;
        const float Is = K.apply_sdw ? cone_T<F>(pyr, pg, K.ss, K.n_steps, p, L) : 0.f;
        const float m = inv_k * (Io * ka + Is * kd);
        src.x *= m; src.y *= m; src.z *= m;
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

// Exact occlusion / shadow and the surrogate's features at given points (texture
// space), for given camera-ray directions and light directions.
template <int F>
__global__ __launch_bounds__(256) void k_cone_points(PyrGeom pg, const void* __restrict__ pyr, ConeK K, int n,
                                                     const float* __restrict__ pos, const float* __restrict__ dirs,
                                                     const float* __restrict__ lights, float* __restrict__ occ,
                                                     float* __restrict__ sdw, float* __restrict__ feat) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const f3 p = ld3(pos + 3 * i);
  const f3 d = norm3(ld3(dirs + 3 * i));
  const f3 L = norm3(ld3(lights + 3 * i));
  f3 k, up, right;
  occ_frame(d, k, up, right);
  if (occ) occ[i] = ct_occlusion<F>(pyr, pg, K, p, k, up, right);
  if (sdw) sdw[i] = cone_T<F>(pyr, pg, K.ss, K.n_steps, p, L);
  if (feat) {
    float f[kNFeat];
    nn_feats<F>(pyr, pg, K.ft, p, k, L, f);
    for (int c = 0; c < kNFeat; ++c) feat[size_t(i) * kNFeat + c] = f[c];
  }
}

// Host side: schedules in double (geometric distances t0 q^j, q = (tmax/t0)^(1/n);
// level log2(r / (r_sigma sigma0)) for the cone radius r at the step midpoint).
static void schedule(int n, double t0, double tmax, double half_deg, double r_sigma, double sigma0, int levels,
                     std::vector<ConeStepF>& out) {
  out.resize(n);
  const double q = std::pow(tmax / t0, 1.0 / n), th = std::tan(half_deg * 3.14159265358979323846 / 180.0);
  for (int j = 0; j < n; ++j) {
    const double a = t0 * std::pow(q, double(j)), b = t0 * std::pow(q, double(j + 1));
    const double tm = 0.5 * (a + b);
    double lod = std::log2(tm * th / (r_sigma * sigma0));
    lod = std::min(std::max(lod, 0.0), double(levels - 1));
    out[j] = ConeStepF{float(tm), float(b - a), float(lod), 0.f};
  }
}

// The surrogate's stencil: kNOcc geometric segments over [t0, occ_max], each read on
// the axis at distance D = sqrt(ab) (level of the axial cone's radius) and at D cos(ring)
// (level of the ring's radius D sin(ring)); kNSdw segments over [t0, sdw_max] along the
// light (level of the shadow cone's radius).  Levels are rounded to whole levels, so
// each input costs one trilinear read.
static void stencil(const cvr_cone_params& P, double occ_max, double sdw_max, int levels, std::vector<StencilF>& st) {
  st.resize(kNStencil);
  const double d2r = 3.14159265358979323846 / 180.0;
  const double tha = std::tan(double(P.occ_half_deg) * d2r), cr = std::cos(double(P.ring_deg) * d2r);
  const double sr = std::sin(double(P.ring_deg) * d2r), ts = std::tan(double(P.sdw_half_deg) * d2r);
  const double t0 = P.t0, rs = double(P.r_sigma) * double(P.sigma0), top = double(levels - 1);
  auto lvl = [&](double r) { return std::min(std::max(std::round(std::log2(r / rs)), 0.0), top); };
  for (int i = 0; i < kNOcc; ++i) {
    const double a = t0 * std::pow(occ_max / t0, double(i) / kNOcc), b = t0 * std::pow(occ_max / t0, double(i + 1) / kNOcc);
    const double D = std::sqrt(a * b);
    st[2 * i] = StencilF{float(D), float(b - a), float(lvl(D * tha)), 0.f};
    st[2 * i + 1] = StencilF{float(D * cr), float(b - a), float(lvl(D * sr)), 0.f};
  }
  for (int i = 0; i < kNSdw; ++i) {
    const double a = t0 * std::pow(sdw_max / t0, double(i) / kNSdw), b = t0 * std::pow(sdw_max / t0, double(i + 1) / kNSdw);
    const double D = std::sqrt(a * b);
    st[2 * kNOcc + i] = StencilF{float(D), float(b - a), float(lvl(D * ts)), 0.f};
  }
}

// A stencil point read at its whole level: the level's geometry copied next to it.
static FeatTap feat_tap(const PyrGeom& pg, const StencilF& s) {
  FeatTap t{};
  const int l = int(s.lod);
  const double log2e = 1.4426950408889634;
  t.dist = s.dist;
  t.seg = float(double(s.seg) * log2e);
  for (int a = 0; a < 3; ++a) {
    t.ts[a] = pg.ts[l][a];
    t.dmax[a] = pg.dmax[l][a];
  }
  t.fall_k = float(log2e / (2.0 * std::ldexp(1.0, 2 * l)));  // log2(e) / (2 (2^l)^2)
  t.nb[0] = pg.nb[l][0];
  t.nb[1] = pg.nb[l][1];
  const size_t bytes = pg.boff[l] * (pg.fmt == 0 ? 32u : 16u);  // records of 32 (fp32) or 16 bytes
  t.boff_lo = unsigned(bytes & 0xFFFFFFFFu);
  t.boff_hi = unsigned(bytes >> 32);
  t.level = l;
  return t;
}

int check_cone_params(const cvr_cone_params& P) {
  if (P.n_steps < 1 || P.n_steps > kMaxSteps || P.n_occ_cones < 1 || P.n_occ_cones > 64 || !(P.t0 > 0) ||
      !(P.r_sigma > 0) || !(P.sigma0 > 0) || !(P.occ_half_deg > 0 && P.occ_half_deg < 90) ||
      !(P.sdw_half_deg > 0 && P.sdw_half_deg < 90) || P.n_levels > kMaxLevels || P.density_precision < 0 ||
      P.density_precision > 2) {
    set_error("cone params: n_steps in [1,256], n_occ_cones in [1,64], t0, r_sigma, sigma0 > 0, half angles in (0,90), "
              "density_precision 0, 1 or 2");
    return -1;
  }
  if (P.apply_occlusion == 0 && P.apply_shadow == 0) {
    set_error("cone params: occlusion or shadow must be enabled (the shading divides by ka+kd)");
    return -1;
  }
  return 0;
}

// Builds the pyramid (if needed) and the per-volume cone tables; fills K's pointers.
static PyrKey cone_pyr_key(const cvr_ctx* c, const cvr_cone_params& P) {
  const int bx = P.base_res[0] > 0 ? P.base_res[0] : c->nx;
  const int by = P.base_res[1] > 0 ? P.base_res[1] : c->ny;
  const int bz = P.base_res[2] > 0 ? P.base_res[2] : c->nz;
  return pyr_key_for(c, bx, by, bz, P.sigma0, P.n_levels, P.density_precision);
}

bool cone_ready(const cvr_ctx* c, const cvr_cone_params& P) {
  return c->d_pyr && c->pyr_valid && c->pyr_key == cone_pyr_key(c, P) && c->d_cone_tab && c->cone_tab_valid &&
         c->cone_tab_pyr == c->pyr_key && std::memcmp(&c->cone_tab_for, &P, sizeof(P)) == 0;
}

int prepare_cone(cvr_ctx* c, const cvr_cone_params& P, hipStream_t st) {
  int rc = check_cone_params(P);
  if (rc) return rc;
  const int bx = P.base_res[0] > 0 ? P.base_res[0] : c->nx;
  const int by = P.base_res[1] > 0 ? P.base_res[1] : c->ny;
  const int bz = P.base_res[2] > 0 ? P.base_res[2] : c->nz;
  rc = build_pyramid(c, bx, by, bz, P.sigma0, P.n_levels, P.density_precision, st);
  if (rc) return rc;
  // the tap table and the step schedules depend on the parameters and the pyramid (its
  // levels, the volume's size) only: rebuilt (a synchronous upload) when one changes
  if (cone_ready(c, P)) return 0;
  c->cone_tab_valid = false;
  const double diag = volume_diagonal(c);
  const double occ_max = P.occ_max_dist > 0 ? P.occ_max_dist : 0.5 * float(diag);
  const double sdw_max = P.sdw_max_dist > 0 ? P.sdw_max_dist : 0.75 * float(diag);
  if (!(occ_max > P.t0) || !(sdw_max > P.t0)) {
    set_error("cone params: march lengths must exceed t0");
    return -1;
  }
  std::vector<ConeStepF> so, ss;
  schedule(P.n_steps, P.t0, occ_max, P.occ_half_deg, P.r_sigma, P.sigma0, c->pyr.levels, so);
  schedule(P.n_steps, P.t0, sdw_max, P.sdw_half_deg, P.r_sigma, P.sigma0, c->pyr.levels, ss);
  std::vector<StencilF> stn;
  stencil(P, occ_max, sdw_max, c->pyr.levels, stn);
  stn.push_back(StencilF{0.f, 1.f, 0.f, 0.f});  // the local opacity: level 0 at the sample
  std::vector<FeatTap> taps(kNFeat);
  for (int i = 0; i < kNFeat; ++i) taps[i] = feat_tap(c->pyr, stn[i]);
  // [taps (kNFeat x 64 B)] [occlusion schedule] [shadow schedule]
  const size_t bytes = kNFeat * sizeof(FeatTap) + (so.size() + ss.size()) * sizeof(ConeStepF);
  if (!c->d_cone_tab) CVR_HIP_CHECK(hipMalloc(&c->d_cone_tab, kNFeat * sizeof(FeatTap) + 2 * kMaxSteps * sizeof(ConeStepF)));
  else CVR_HIP_CHECK(hipDeviceSynchronize());  // earlier renders may still read the old table
  std::vector<char> buf(bytes);
  std::memcpy(buf.data(), taps.data(), kNFeat * sizeof(FeatTap));
  std::memcpy(buf.data() + kNFeat * sizeof(FeatTap), so.data(), so.size() * sizeof(ConeStepF));
  std::memcpy(buf.data() + kNFeat * sizeof(FeatTap) + so.size() * sizeof(ConeStepF), ss.data(), ss.size() * sizeof(ConeStepF));
  CVR_HIP_CHECK(hipMemcpy(c->d_cone_tab, buf.data(), bytes, hipMemcpyHostToDevice));
  c->cone_n_steps = P.n_steps;
  c->cone_tab_for = P;
  c->cone_tab_pyr = c->pyr_key;
  c->cone_tab_valid = true;
  return 0;
}

ConeK cone_constants(const cvr_ctx* c, const cvr_cone_params& P, const float L[3]) {
  ConeK K{};
  K.n_steps = P.n_steps;
  K.n_occ = P.n_occ_cones;
  K.apply_occ = P.apply_occlusion;
  K.apply_sdw = P.apply_shadow;
  const double pi = 3.14159265358979323846;
  K.ring_cos = float(std::cos(double(P.ring_deg) * pi / 180.0));
  K.ring_sin = float(std::sin(double(P.ring_deg) * pi / 180.0));
  K.ka = P.ka;
  K.kd = P.kd;
  double l = std::sqrt(double(L[0]) * L[0] + double(L[1]) * L[1] + double(L[2]) * L[2]);
  for (int i = 0; i < 3; ++i) K.L[i] = float(L[i] / l);
  K.ft = static_cast<const FeatTap*>(c->d_cone_tab);
  K.so = reinterpret_cast<const ConeStepF*>(K.ft + kNFeat);
  K.ss = K.so + P.n_steps;
  return K;
}

void light_dir(const RayFrame& rf, const cvr_light* light, float L[3]) {
  for (int i = 0; i < 3; ++i) L[i] = light ? light->forward[i] : -rf.f[i];
}

int launch_cone(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st) {
  const cvr_cone_params& P = c->cone;
  int rc = prepare_cone(c, P, st);
  if (rc) return rc;
  float L[3];
  light_dir(rf, light, L);
  ConeK K = cone_constants(c, P, L);
  VolGeom g = geom_of(c);
  const float step = P.step_size > 0 ? P.step_size : default_step(c);
  const int tiles = ((rf.W + 15) / 16) * ((rf.H + 15) / 16);
  const size_t shm = size_t(c->tf_size) * sizeof(float4);
  if (shm > 64 * 1024) { set_error("cone renderer: transfer function larger than 4096 entries"); return -4; }
  unsigned long long* cnt = c->collect_stats ? c->d_counters : nullptr;
  float4* o4 = reinterpret_cast<float4*>(out);
  with_types(c->dtype, c->pyr.fmt, [&](auto t, auto f) {
    using T = typename decltype(t)::type;
    hipLaunchKernelGGL((k_cone_render<T, decltype(f)::value>), dim3(tiles), dim3(256), shm, st, rf, g, c->pyr,
                       c->d_pyr_oct, K, c->d_tf_rgbt, c->tf_size, step, P.et_threshold, o4, cnt);
  });
  CVR_HIP_CHECK(hipGetLastError());
  return 0;
}

int cone_points(cvr_ctx* c, const cvr_cone_params& P, int n, const float* d_pos, const float* d_dirs,
                const float* d_lights, float* d_occ, float* d_sdw, float* d_feat, hipStream_t st) {
  int rc = prepare_cone(c, P, st);
  if (rc) return rc;
  float L[3] = {0, 0, 1};
  ConeK K = cone_constants(c, P, L);
  with_fmt(c->pyr.fmt, [&](auto f) {
    hipLaunchKernelGGL((k_cone_points<decltype(f)::value>), dim3((n + 255) / 256), dim3(256), 0, st, c->pyr,
                       c->d_pyr_oct, K, n, d_pos, d_dirs, d_lights, d_occ, d_sdw, d_feat);
  });
  CVR_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace cvr
