// Plain single-pass ray casting (front-to-back emission-absorption), gfx950.
//
// Semantics follow cppvolrend/structured/rc1pass/ray_marching_1p.comp:156-250
// (USE_TRANSPARENCY off): midpoint samples of the segment [s, s+h], h = min(step,
// D - s); the TF holds (r, g, b, extinction); alpha = 1 - exp(-ext*h); composite
// front to back; stop once accumulated alpha exceeds the threshold.
//
// With gradient shading on, every opaque sample's colour goes through ShadeBlinnPhong
// (:48-81) before it is composited: rgb*(ka + kd*max(0, n.l)) + specular*ks*
// max(0, h.n)^shininess, n the normalised GL_LINEAR gradient (samples whose gradient is
// the zero vector keep their colour).
//
// Mapping: one lane per pixel, a wavefront covers an 8x8 pixel tile, a 256-lane
// workgroup a 16x16 tile; workgroups are re-ordered so each XCD marches one
// contiguous band of tiles.  The 1D transfer function is staged in LDS.  The
// sample position is recomputed from the sample index each step (no running
// sum), so the step schedule does not drift in fp32.  Positions do not depend on the
// compositing, so the march is software-pipelined: the corner-block loads of the next
// kGroup samples are in flight while the current kGroup are composited, and a group's
// transfer-function reads (LDS) are all issued before its first exit test (128^3 u8 at
// 512^2: 0.155 ms one sample at a time, 0.106 ms pipelined, 0.094 ms with the reads
// hoisted).
#include "cvr_ctx.hpp"
#include <type_traits>

namespace cvr {

constexpr int kGroup = 4;  // samples whose volume loads are issued together
template <class T> struct VoxTag { typedef T type; };

template <typename T, bool TF_LDS, bool PH>
__global__ __launch_bounds__(256) void k_rc1p(RayFrame rf, VolGeom g, const float4* __restrict__ tf, int ntf,
                                               float step, float et, PhongK ph, const float4* __restrict__ grad,
                                               float ka, float kd, float4* __restrict__ out,
                                               unsigned long long* __restrict__ counters) {
  extern __shared__ float4 s_tf[];
  if (TF_LDS) {
    for (int i = threadIdx.x; i < ntf; i += blockDim.x) s_tf[i] = tf[i];
    __syncthreads();
  }
  const float4* tfp = TF_LDS ? s_tf : tf;
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tiles_x = (rf.W + 15) >> 4;
  int tx, ty;
  super_tile(bid, tiles_x, (rf.H + 15) >> 4, 16, 32, tx, ty);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = tx * 16 + (w & 1) * 8 + (lane & 7);
  const int py = ty * 16 + (w >> 1) * 8 + (lane >> 3);
  if (px >= rf.W || py >= rf.H) return;

  f3 dir = camera_ray(rf, px, py);
  f3 eye = ld3(rf.eye);
  float tn, tf_;
  float4 dst = make_float4(0.f, 0.f, 0.f, 0.f);
  unsigned int nsamp = 0, nshade = 0;
  if (ray_box(eye, dir, g.G, tn, tf_)) {
    const float D = fabsf(tf_ - tn);
    const f3 tp = eye + dir * tn + mk3(0.5f * g.G[0], 0.5f * g.G[1], 0.5f * g.G[2]);
    typedef typename VolOct<T>::type RT;
    // sample j sits at the midpoint of [s, s + h], s = j step, h = min(step, D - s); for
    // j past the end the position is still finite and its clamped block in bounds
    auto pos = [&](int j) {
      const float s = float(j) * step;
      return tp + dir * (s + 0.5f * fminf(step, D - s));
    };
    // Samples in groups of kGroup; while one group is composited the loads of the next
    // are in flight (two register sets, a and b, alternate).  Returns true when the ray
    // ends inside the group.
    auto group = [&](const RT* rec, int j0) {
      // the group's densities and transfer-function entries first (no side effects), so
      // the table reads of later samples need not wait for the earlier samples' exit tests
      float4 srcs[kGroup];
#pragma unroll
      for (int k = 0; k < kGroup; ++k) {
        const int j = j0 + k;
        const float s = float(j) * step;
        const float h = fminf(step, D - s);
        const f3 p = tp + dir * (s + 0.5f * h);
        srcs[k] = sample_tf(tfp, ntf, volume_oct_lerp<T>(g, rec[k], p));
      }
#pragma unroll
      for (int k = 0; k < kGroup; ++k) {
        const int j = j0 + k;
        const float s = float(j) * step;
        if (!(s < D) || j >= kMaxSamples) return true;
        const float h = fminf(step, D - s);
        const f3 p = tp + dir * (s + 0.5f * h);
        float4 src = srcs[k];
        ++nsamp;
        if (src.w > 0.0f) {
          ++nshade;
          float dd, sp;
          if (PH && phong_terms(grad, g, ph, p, eye, dd, sp)) {
            const float m = ka + kd * dd, s2 = ph.ks * sp;
            src.x = src.x * m + ph.spec[0] * s2;
            src.y = src.y * m + ph.spec[1] * s2;
            src.z = src.z * m + ph.spec[2] * s2;
          }
          const float a = 1.0f - __expf(-src.w * h);
          const float wgt = (1.0f - dst.w) * a;
          dst.x += wgt * src.x;
          dst.y += wgt * src.y;
          dst.z += wgt * src.z;
          dst.w += wgt;
          if (dst.w > et) return true;
        }
      }
      return false;
    };
    RT a[kGroup], b[kGroup];
#pragma unroll
    for (int k = 0; k < kGroup; ++k) a[k] = volume_oct_load<T>(g, pos(k));
    for (int j0 = 0;; j0 += 2 * kGroup) {
#pragma unroll
      for (int k = 0; k < kGroup; ++k) b[k] = volume_oct_load<T>(g, pos(j0 + kGroup + k));
      if (group(a, j0)) break;
#pragma unroll
      for (int k = 0; k < kGroup; ++k) a[k] = volume_oct_load<T>(g, pos(j0 + 2 * kGroup + k));
      if (group(b, j0 + kGroup)) break;
    }
  }
  out[size_t(py) * rf.W + px] = dst;
  if (counters) {
    atomicAdd(&counters[0], (unsigned long long)nsamp);
    atomicAdd(&counters[1], (unsigned long long)nshade);
  }
}

int phong_setup(cvr_ctx* c, int on, int gradient_type, float ks, float shininess, const float spec[3],
                const RayFrame& rf, const cvr_light* light, PhongK& ph, hipStream_t st) {
  ph = PhongK{};
  if (!on) return 0;
  if (!(shininess > 0.f)) {
    set_error("cvr_render: Blinn-Phong shininess must be > 0");
    return CVR_ERR_INVALID;
  }
  if (const int rc = build_gradient(c, gradient_type, st)) return rc;
  ph.on = 1;
  ph.ks = ks;
  ph.shininess = shininess;
  for (int i = 0; i < 3; ++i) {
    ph.spec[i] = spec[i];
    // without a light: the camera's own light at the eye (renderingmanager.cpp:421-426)
    ph.lpos[i] = light ? light->position[i] : rf.eye[i];
  }
  return 0;
}

int launch_rc1p(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st) {
  const cvr_rc1p_params& P = c->rc1p;
  VolGeom g = geom_of(c);
  float step = P.step_size > 0 ? P.step_size : default_step(c);
  float et = P.et_threshold;
  PhongK ph;
  if (const int rc = phong_setup(c, P.apply_gradient_shading, P.gradient_type, P.ks, P.shininess, P.specular, rf, light,
                                 ph, st))
    return rc;
  const float4* gr = ph.on ? c->d_grad : nullptr;
  int tiles = ((rf.W + 15) / 16) * ((rf.H + 15) / 16);
  bool lds = c->tf_size <= 4096;
  size_t shm = lds ? size_t(c->tf_size) * sizeof(float4) : 0;
  unsigned long long* cnt = c->collect_stats ? c->d_counters : nullptr;
  float4* o4 = reinterpret_cast<float4*>(out);
  // the Blinn-Phong branch is a compile-time variant: without it the unrolled groups of
  // the pipelined march stay small enough for twice the waves per SIMD
  auto launch = [&](auto t, auto l, auto p) {
    using T = typename decltype(t)::type;
    constexpr bool L = decltype(l)::value, PHv = decltype(p)::value;
    hipLaunchKernelGGL((k_rc1p<T, L, PHv>), dim3(tiles), dim3(256), L ? shm : 0, st, rf, g, c->d_tf_rgbt, c->tf_size,
                       step, et, ph, gr, P.ka, P.kd, o4, cnt);
  };
  auto by_phong = [&](auto t, auto l) {
    if (ph.on) launch(t, l, std::true_type{});
    else launch(t, l, std::false_type{});
  };
  auto by_lds = [&](auto t) {
    if (lds) by_phong(t, std::true_type{});
    else by_phong(t, std::false_type{});
  };
  if (c->dtype == CVR_U8) by_lds(VoxTag<uint8_t>{});
  else by_lds(VoxTag<uint16_t>{});
  CVR_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace cvr
