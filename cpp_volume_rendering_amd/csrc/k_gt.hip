// Ray-sampled cone ground truth (gfx950): the renderer of RC1PConeLightGroundTruthSteps.
//
// Each opaque sample of the camera march (rc1pcrtgt/gt_ray_marching.comp:371-477)
// casts a fixed set of rays inside the occlusion cone around the direction back to the
// eye and inside the shadow cone towards the light (:109-252).  A ray starts
// `initial_gap` from the sample, advances in steps of `ray_step`, stops before leaving
// the volume box or once its opacity passes 0.99, and accumulates transmittance with
// the trapezoid rule over the extinction of the RGBt transfer function.  Rays are
// weighted by the cosine to the cone axis.  The reference runs the camera march one
// sample per dispatch with the state in 16-bit images (crtgtrenderer.cpp:262-325);
// here every ray finishes in one dispatch in fp32.
//
// Ray direction tables come from a default-seeded std::default_random_engine, theta
// uniform over the aperture and phi over the full turn (crtgtrenderer.cpp:131-185).
#include "cvr_ctx.hpp"
#include <cmath>
#include <random>
#include <vector>

namespace cvr {

struct GtK {
  const float* occ_rays;  // (x, y, z) per ray in the (right, up, axis) frame
  const float* sdw_rays;
  int n_occ, n_sdw, apply_occ, apply_sdw, shadow_type;
  float occ_dist, sdw_dist, gap, rstep, ka, kd;
  float lpos[3], lfwd[3], lup[3], lright[3];  // lfwd: the light-camera forward the shader receives
};

// Transmittance of one ray (ConeOcclusionEvaluationRayCasting inner loop, :125-159).
template <typename T>
__device__ __forceinline__ float gt_ray(const VolGeom& g, const float4* __restrict__ tf, int ntf, f3 p, f3 w,
                                        float dist, float gap, float rstep) {
  float Vt = 1.0f;
  float st0 = sample_tf(tf, ntf, sample_volume_oct<T>(g, p + w * gap)).w;
  for (int j = 0; j < kMaxSamples; ++j) {
    const float s = gap + float(j) * rstep;
    if (!(s < dist)) break;
    const float h = fminf(rstep, dist - s);
    const f3 at = p + w * (s + h);
    if (at.x < 0.f || at.x > g.G[0] || at.y < 0.f || at.y > g.G[1] || at.z < 0.f || at.z > g.G[2]) break;
    const float st1 = sample_tf(tf, ntf, sample_volume_oct<T>(g, at)).w;
    Vt *= __expf(-((st0 + st1) * 0.5f) * h);
    if ((1.0f - Vt) > 0.99f) break;
    st0 = st1;
  }
  return Vt;
}

// Cosine-weighted mean transmittance of a ray set in the frame (right, up, axis).
template <typename T>
__device__ __forceinline__ float gt_cone(const VolGeom& g, const float4* __restrict__ tf, int ntf,
                                         const float* __restrict__ rays, int n, f3 p, f3 right, f3 up, f3 axis,
                                         float dist, float gap, float rstep) {
  float S = 0.f, W = 0.f;
  for (int r = 0; r < n; ++r) {
    const f3 w = norm3(right * rays[3 * r] + up * rays[3 * r + 1] + axis * rays[3 * r + 2]);
    const float Vt = gt_ray<T>(g, tf, ntf, p, w, dist, gap, rstep);
    const float rw = dot3(axis, w);
    S += Vt * rw;
    W += rw;
  }
  return S / W;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gt_render(RayFrame rf, VolGeom g, GtK K, const float4* __restrict__ tf,
                                                   int ntf, float step, float et, float4* __restrict__ out,
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
    const f3 vdir = norm3(mk3(-dir.x, -dir.y, -dir.z));
    const f3 half = mk3(0.5f * g.G[0], 0.5f * g.G[1], 0.5f * g.G[2]);
    const float D = tf_ - tn;
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
        float Io = 0.f, Is = 0.f;
        if (K.apply_occ)
          Io = gt_cone<T>(g, s_tf, ntf, K.occ_rays, K.n_occ, p, right, up, vdir, K.occ_dist, K.gap, K.rstep);
        if (K.apply_sdw && K.shadow_type != 1) {
          f3 sd, su, sr;
          if (K.shadow_type == 2) {
            sd = ld3(K.lfwd);
            su = ld3(K.lup);
            sr = ld3(K.lright);
          } else {
            sd = norm3(ld3(K.lpos) - (p - half));
            su = norm3(cross3(sd, ld3(K.lright)));
            sr = norm3(cross3(sd, su));
          }
          Is = gt_cone<T>(g, s_tf, ntf, K.sdw_rays, K.n_sdw, p, sr, su, sd, K.sdw_dist, K.gap, K.rstep);
        }
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

// ------------------------------------------------------------------ host
// Ray set of one cone (crtgtrenderer.cpp:139-151): theta and phi in float as the
// reference draws them; the two rotations (about y by theta, about the axis by phi)
// in double.
static void draw_rays(std::default_random_engine& gen, int n, float aperture_deg, std::vector<float>& out) {
  std::uniform_real_distribution<float> U(0.0f, 1.0f);
  const float pi = 3.14159265358979323846f;
  out.resize(size_t(n) * 3);
  for (int i = 0; i < n; ++i) {
    const float theta = U(gen) * (pi * (aperture_deg / 180.f));
    const float phi = U(gen) * (pi * (360.f / 180.f));
    // rotate (0,0,1) about y by theta, then about z by phi
    const double st = std::sin(double(theta)), ct = std::cos(double(theta));
    const double sp = std::sin(double(phi)), cp = std::cos(double(phi));
    double v[3] = {st * cp, st * sp, ct};
    const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int c = 0; c < 3; ++c) out[size_t(i) * 3 + c] = float(v[c] / l);
  }
}

int check_gt_params(const cvr_gt_params& P) {
  if (P.apply_occlusion == 0 && P.apply_shadow == 0) {
    set_error("gt params: occlusion or shadow must be enabled (the shading divides by ka+kd)");
    return CVR_ERR_INVALID;
  }
  if (P.occ_rays < 1 || P.occ_rays > 4096 || P.sdw_rays < 1 || P.sdw_rays > 4096 || !(P.ray_step > 0) ||
      !(P.initial_gap >= 0) || P.shadow_type < 0 || P.shadow_type > 2 || !(P.et_threshold > 0) ||
      !(P.occ_aperture_deg >= 0) || !(P.sdw_aperture_deg >= 0)) {
    set_error("gt params: ray counts in [1,4096], ray_step > 0, initial_gap >= 0, shadow_type in {0,1,2}, apertures >= 0");
    return CVR_ERR_INVALID;
  }
  return CVR_OK;
}

int prepare_gt(cvr_ctx* c, const cvr_gt_params& P) {
  int rc = check_gt_params(P);
  if (rc) return rc;
  const bool same = c->d_gt_rays && c->gt_rays_for.occ_rays == P.occ_rays && c->gt_rays_for.sdw_rays == P.sdw_rays &&
                    c->gt_rays_for.occ_aperture_deg == P.occ_aperture_deg &&
                    c->gt_rays_for.sdw_aperture_deg == P.sdw_aperture_deg;
  if (same) return CVR_OK;
  std::default_random_engine gen;  // default seed, one stream: occlusion rays then shadow rays
  draw_rays(gen, P.occ_rays, P.occ_aperture_deg, c->gt_occ);
  draw_rays(gen, P.sdw_rays, P.sdw_aperture_deg, c->gt_sdw);
  if (c->d_gt_rays) CVR_HIP_CHECK(hipFree(c->d_gt_rays));
  c->d_gt_rays = nullptr;
  const size_t n = c->gt_occ.size() + c->gt_sdw.size();
  if (hipMalloc(&c->d_gt_rays, n * sizeof(float)) != hipSuccess) {
    c->d_gt_rays = nullptr;
    set_error("gt: out of device memory");
    return CVR_ERR_NOMEM;
  }
  std::vector<float> all(c->gt_occ);
  all.insert(all.end(), c->gt_sdw.begin(), c->gt_sdw.end());
  CVR_HIP_CHECK(hipMemcpy(c->d_gt_rays, all.data(), n * sizeof(float), hipMemcpyHostToDevice));
  c->gt_rays_for = P;
  return CVR_OK;
}

int launch_gt(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st) {
  const cvr_gt_params& P = c->gt;
  int rc = prepare_gt(c, P);
  if (rc) return rc;
  GtK K{};
  K.occ_rays = c->d_gt_rays;
  K.sdw_rays = c->d_gt_rays + c->gt_occ.size();
  K.n_occ = P.occ_rays;
  K.n_sdw = P.sdw_rays;
  K.apply_occ = P.apply_occlusion;
  K.apply_sdw = P.apply_shadow;
  K.shadow_type = P.shadow_type;
  // m_occ/sdw_cone_distance_eval = diagonal * 0.50f / 0.75f (crtgtrenderer.cpp:110-114)
  const float diag = volume_diagonal(c);
  K.occ_dist = P.occ_max_dist > 0 ? P.occ_max_dist : diag * 0.50f;
  K.sdw_dist = P.sdw_max_dist > 0 ? P.sdw_max_dist : diag * 0.75f;
  K.gap = P.initial_gap;
  K.rstep = P.ray_step;
  K.ka = P.ka;
  K.kd = P.kd;
  for (int i = 0; i < 3; ++i) {
    // the shader's LightCamForward is minus the light camera's forward (crtgtrenderer.cpp:224-234)
    K.lpos[i] = light ? light->position[i] : rf.eye[i];
    K.lfwd[i] = -(light ? light->forward[i] : -rf.f[i]);
    K.lup[i] = light ? light->up[i] : rf.u[i];
    K.lright[i] = light ? light->right[i] : rf.s[i];
  }
  VolGeom g = geom_of(c);
  const float step = P.step_size > 0 ? P.step_size : default_step(c);
  const int tiles = ((rf.W + 15) / 16) * ((rf.H + 15) / 16);
  const size_t shm = size_t(c->tf_size) * sizeof(float4);
  if (shm > 64 * 1024) { set_error("gt renderer: transfer function larger than 4096 entries"); return CVR_ERR_UNSUPPORTED; }
  unsigned long long* cnt = c->collect_stats ? c->d_counters : nullptr;
  if (c->dtype == CVR_U8)
    hipLaunchKernelGGL(k_gt_render<uint8_t>, dim3(tiles), dim3(256), shm, st, rf, g, K, c->d_tf_rgbt, c->tf_size, step,
                       P.et_threshold, reinterpret_cast<float4*>(out), cnt);
  else
    hipLaunchKernelGGL(k_gt_render<uint16_t>, dim3(tiles), dim3(256), shm, st, rf, g, K, c->d_tf_rgbt, c->tf_size,
                       step, P.et_threshold, reinterpret_cast<float4*>(out), cnt);
  CVR_HIP_CHECK(hipGetLastError());
  return CVR_OK;
}

}  // namespace cvr
