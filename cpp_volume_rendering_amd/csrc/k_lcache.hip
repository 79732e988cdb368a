// Object-space light cache of the EBS and DOS renderers (gfx950): the march that reads
// it and its buffer.  The cache kernels themselves live with their techniques
// (k_ebs.hip: k_lc_ebs, k_dos.hip: k_lc_dos).
//
// PreIlluminationStructuredVolume (cppvolrend/utils/preillumination.cpp:7-80) keeps a
// 2-channel (occlusion, shadow) volume, 32^3 by default, rebuilt every Update() when
// active (ebsrenderer.cpp:127-134, dosrcrenderer.cpp:134-143).  The frame is then marched
// by _common_shaders/obj_ray_marching.comp:223-333: every opaque sample reads the cache
// with GL_LINEAR / CLAMP_TO_EDGE at tex_pos / VolumeScaledSizes and shades
// rgb <- (rgb*Iocc*ka + rgb*Isdw*kd) / (ka + kd), or with Phong shading on
// (rgb*Iocc*ka + Isdw*rgb*kd*max(0, n.l)) / (ka + kd) + Isdw*ks*specular*max(0, h.n)^n
// (:186-211; a sample whose gradient vanishes keeps its colour).  Here the
// cache is fp32 (the reference stores RG16F), laid out x fastest as float2; at the
// default size it is 256 KiB and stays in L2, so its 8 taps per sample are cache hits.
#include "cvr_ctx.hpp"
#include <type_traits>

namespace cvr {

constexpr int kMaxLightCacheRes = 512;
constexpr int kLcGroup = 4;  // samples whose volume loads are issued together
template <class T> struct LcVox { typedef T type; };

// GL_LINEAR of the cache at texel coordinates p * tex_mul - 1/2 (clamped to the edge).
__device__ __forceinline__ float2 lc_sample(const float2* __restrict__ lc, int rx, int ry, int rz, const float tm[3],
                                            f3 p) {
  int x0, x1, y0, y1, z0, z1;
  float fx, fy, fz;
  lin_axis(p.x * tm[0] - 0.5f, rx, x0, x1, fx);
  lin_axis(p.y * tm[1] - 0.5f, ry, y0, y1, fy);
  lin_axis(p.z * tm[2] - 0.5f, rz, z0, z1, fz);
  const size_t sy = size_t(rx), sz = size_t(rx) * size_t(ry);
  const size_t b00 = size_t(y0) * sy + size_t(z0) * sz, b10 = size_t(y1) * sy + size_t(z0) * sz;
  const size_t b01 = size_t(y0) * sy + size_t(z1) * sz, b11 = size_t(y1) * sy + size_t(z1) * sz;
  const float2 c000 = lc[b00 + x0], c100 = lc[b00 + x1], c010 = lc[b10 + x0], c110 = lc[b10 + x1];
  const float2 c001 = lc[b01 + x0], c101 = lc[b01 + x1], c011 = lc[b11 + x0], c111 = lc[b11 + x1];
  auto lerp2 = [](float2 a, float2 b, float t) { return make_float2(a.x + t * (b.x - a.x), a.y + t * (b.y - a.y)); };
  const float2 c0 = lerp2(lerp2(c000, c100, fx), lerp2(c010, c110, fx), fy);
  const float2 c1 = lerp2(lerp2(c001, c101, fx), lerp2(c011, c111, fx), fy);
  return lerp2(c0, c1, fz);
}

struct LcK {
  int rx, ry, rz;
  float tex_mul[3];  // cache texels per world unit: res / G
  float ka, kd;      // 0 for a term that is off
  float ks;          // Phong specular weight, 0 without shadows (Kspecular counts with ApplyShadow)
};

// PH: the Blinn-Phong branch, a compile-time variant so that the unrolled sample groups
// without it stay small.
template <typename T, bool PH>
__global__ __launch_bounds__(256) void k_lc_render(RayFrame rf, VolGeom g, LcK L, const float2* __restrict__ lc,
                                                   const float4* __restrict__ tf, int ntf, float step, float et,
                                                   PhongK ph, const float4* __restrict__ grad, float4* __restrict__ out,
                                                   unsigned long long* __restrict__ counters) {
  extern __shared__ float4 s_tf[];
  for (int i = threadIdx.x; i < ntf; i += blockDim.x) s_tf[i] = tf[i];
  __syncthreads();
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tiles_x = (rf.W + 15) >> 4;
  int tx, ty;
  super_tile(bid, tiles_x, (rf.H + 15) >> 4, 16, 32, tx, ty);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = tx * 16 + (w & 1) * 8 + (lane & 7);
  const int py = ty * 16 + (w >> 1) * 8 + (lane >> 3);
  if (px >= rf.W || py >= rf.H) return;
  const f3 dir = camera_ray(rf, px, py);
  const f3 eye = ld3(rf.eye);
  float tn, tf_;
  float4 dst = make_float4(0.f, 0.f, 0.f, 0.f);
  unsigned int nsamp = 0, nshade = 0;
  const float inv_k = 1.0f / (L.ka + L.kd);
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
    // Samples in groups of kLcGroup, as the plain march: the corner blocks of the next
    // group in flight while one is composited (two register sets, a and b), and a group's
    // transfer-function and cache reads issued before its first exit test.  Returns true
    // when the ray ends inside the group.
    auto group = [&](const RT* rec, int j0) {
      float4 srcs[kLcGroup];
      float2 Is[kLcGroup];
#pragma unroll
      for (int k = 0; k < kLcGroup; ++k) {
        const f3 p = pos(j0 + k);
        srcs[k] = sample_tf(s_tf, ntf, volume_oct_lerp<T>(g, rec[k], p));
        Is[k] = make_float2(0.f, 0.f);
        if (srcs[k].w > 0.0f) Is[k] = lc_sample(lc, L.rx, L.ry, L.rz, L.tex_mul, p);
      }
#pragma unroll
      for (int k = 0; k < kLcGroup; ++k) {
        const int j = j0 + k;
        const float s = float(j) * step;
        if (!(s < D) || j >= kMaxSamples) return true;
        const float h = fminf(step, D - s);
        float4 src = srcs[k];
        ++nsamp;
        if (src.w > 0.0f) {
          ++nshade;
          const float2 I = Is[k];
          if (PH) {
            const f3 p = tp + dir * (s + 0.5f * h);
            float dd, sp;
            if (phong_terms(grad, g, ph, p, eye, dd, sp)) {
              const float s2 = I.y * (L.ks * sp);
              src.x = inv_k * (src.x * I.x * L.ka + I.y * (src.x * L.kd * dd)) + s2 * ph.spec[0];
              src.y = inv_k * (src.y * I.x * L.ka + I.y * (src.y * L.kd * dd)) + s2 * ph.spec[1];
              src.z = inv_k * (src.z * I.x * L.ka + I.y * (src.z * L.kd * dd)) + s2 * ph.spec[2];
            }
          } else {
            const float m = inv_k * (I.x * L.ka + I.y * L.kd);
            src.x *= m;
            src.y *= m;
            src.z *= m;
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
    RT ra[kLcGroup], rb[kLcGroup];
#pragma unroll
    for (int k = 0; k < kLcGroup; ++k) ra[k] = volume_oct_load<T>(g, pos(k));
    for (int j0 = 0;; j0 += 2 * kLcGroup) {
#pragma unroll
      for (int k = 0; k < kLcGroup; ++k) rb[k] = volume_oct_load<T>(g, pos(j0 + kLcGroup + k));
      if (group(ra, j0)) break;
#pragma unroll
      for (int k = 0; k < kLcGroup; ++k) ra[k] = volume_oct_load<T>(g, pos(j0 + 2 * kLcGroup + k));
      if (group(rb, j0 + kLcGroup)) break;
    }
  }
  out[size_t(py) * rf.W + px] = dst;
  if (counters) {
    atomicAdd(&counters[0], (unsigned long long)nsamp);
    atomicAdd(&counters[1], (unsigned long long)nshade);
  }
}

int lc_prepare(cvr_ctx* c, const int res[3]) {
  for (int i = 0; i < 3; ++i)
    if (res[i] < 1 || res[i] > kMaxLightCacheRes) {
      set_error("light cache: light_cache_res must be in [1, 512] per axis");
      return CVR_ERR_INVALID;
    }
  const size_t n = size_t(res[0]) * size_t(res[1]) * size_t(res[2]);
  if (n > c->lc_cap) {
    if (c->d_lc) CVR_HIP_CHECK(hipFree(c->d_lc));
    c->d_lc = nullptr;
    c->lc_cap = 0;
    CVR_HIP_CHECK(hipMalloc(&c->d_lc, n * sizeof(float2)));
    c->lc_cap = n;
  }
  for (int i = 0; i < 3; ++i) c->lc_res[i] = res[i];
  c->lc_valid = false;
  return CVR_OK;
}

int launch_lc_render(cvr_ctx* c, const RayFrame& rf, int apply_occ, int apply_sdw, float ka, float kd, float step,
                     float et, const PhongK& ph, float* out, hipStream_t st) {
  LcK L{};
  const VolGeom g = geom_of(c);
  L.rx = c->lc_res[0];
  L.ry = c->lc_res[1];
  L.rz = c->lc_res[2];
  const int res[3] = {L.rx, L.ry, L.rz};
  for (int i = 0; i < 3; ++i) L.tex_mul[i] = float(double(res[i]) / double(g.G[i]));
  L.ka = apply_occ ? ka : 0.f;
  L.kd = apply_sdw ? kd : 0.f;
  L.ks = apply_sdw ? ph.ks : 0.f;
  const int tiles = ((rf.W + 15) / 16) * ((rf.H + 15) / 16);
  const size_t shm = size_t(c->tf_size) * sizeof(float4);
  if (shm > 64 * 1024) {
    set_error("cvr_render: light-cache march: transfer function larger than 4096 entries (staged in LDS)");
    return -4;
  }
  unsigned long long* cnt = c->collect_stats ? c->d_counters : nullptr;
  auto launch = [&](auto t, auto p) {
    using T = typename decltype(t)::type;
    constexpr bool PHv = decltype(p)::value;
    hipLaunchKernelGGL((k_lc_render<T, PHv>), dim3(tiles), dim3(256), shm, st, rf, g, L, c->d_lc, c->d_tf_rgbt,
                       c->tf_size, step, et, ph, c->d_grad, reinterpret_cast<float4*>(out), cnt);
  };
  auto by_phong = [&](auto t) {
    if (ph.on) launch(t, std::true_type{});
    else launch(t, std::false_type{});
  };
  if (c->dtype == CVR_U8) by_phong(LcVox<uint8_t>{});
  else by_phong(LcVox<uint16_t>{});
  CVR_HIP_CHECK(hipGetLastError());
  c->lc_valid = true;
  return CVR_OK;
}

}  // namespace cvr
