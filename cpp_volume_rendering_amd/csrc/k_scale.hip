// Multi-scale output (gfx950): the rendered frame resampled to the screen size, as
// RenderFrameToScreen does it (libs/vis_utils/renderoutputframe.cpp:243-535, shaders
// vis_utils/shader/renderoutputframe/*.comp).  Three modes:
//  - multisample: the frame rendered at a multiple of the screen size, read with
//    GL_LINEAR / CLAMP_TO_EDGE at each screen pixel centre (multisample_filter.comp);
//  - downscale: the same, filtered: out = s_r s_c sum_ij k((x_r - (i+1/2)/n_r) T_r)
//    k((x_c - (j+1/2)/n_c) T_c) f_ij over the kernel's support (downscaling_filter.comp),
//    then, for the cardinal kernels, their digital filter on the result;
//  - upscale: the frame rendered at a fraction of the screen size; for the cardinal
//    kernels its digital filter first (on a copy here: the reference filters the
//    rendered frame in place), then out = sum_ij k(xi_r - i) k(xi_c - j) f_ij with
//    xi = x n - 1/2 (upscaling_filter.comp).
// Kernels (box_filter.comp ... cardinalomoms_filter.comp; Nehab and Hoppe, "A fresh look
// at generalized sampling", 2014): box, hat, Catmull-Rom, Mitchell-Netravali (B = C =
// 1/3), cardinal cubic B-spline (scaled by 6) and cardinal O-MOMS3 (scaled by 5.25).
// The digital filter inverts the kernel's sampled convolution (1, a, 1) with a = 4
// (B-spline) or 3.25 (O-MOMS): forward and backward substitution of its LU factors
// L_0 = 1/(a+1), L_i = 1/(a - L_{i-1}), truncated once they settle (8 resp. 9 factors),
// as vis_utils/filters/utils.hpp:17-44 does along every row and then every column.
// texelFetch outside the frame reads 0 here (GL leaves it undefined).
#include "cvr_ctx.hpp"

#include <cmath>

namespace cvr {

__device__ __forceinline__ float cubic_piece(int k, float u, int which) {
  // which: 0 -> k0 (x in (1, 2], u = 2 - x), 1 -> k1 (x in [0, 1], u = 1 - x)
  switch (k) {
    case CVR_FILTER_CATMULL_ROM:
      return which == 0 ? ((0.5f * u - 0.5f) * u) * u : ((-1.5f * u + 2.0f) * u + 0.5f) * u;
    case CVR_FILTER_MITCHELL_NETRAVALI:
      return which == 0 ? (((7.0f / 18.0f) * u - 1.0f / 3.0f) * u) * u
                        : (((-7.0f / 6.0f) * u + 1.5f) * u + 0.5f) * u + 1.0f / 18.0f;
    case CVR_FILTER_CARDINAL_BSPLINE3:
      return which == 0 ? (u * u) * u : ((-3.0f * u + 3.0f) * u + 3.0f) * u + 1.0f;
    default:  // O-MOMS3
      return which == 0 ? ((0.875f * u) * u + 0.125f) * u : ((-2.625f * u + 2.625f) * u + 2.25f) * u + 1.0f;
  }
}

__device__ __forceinline__ float kernel_support(int k) {
  return k == CVR_FILTER_BOX ? 1.f : (k == CVR_FILTER_HAT ? 2.f : 4.f);
}

__device__ __forceinline__ float kernel_weight(int k, float x) {
  if (k == CVR_FILTER_BOX) return (x <= -0.5f || x > 0.5f) ? 0.f : 1.f;
  x = fabsf(x);
  if (k == CVR_FILTER_HAT) return x > 1.f ? 0.f : 1.f - x;
  return x > 2.f ? 0.f : (x > 1.f ? cubic_piece(k, 2.f - x, 0) : cubic_piece(k, 1.f - x, 1));
}

__device__ __forceinline__ float4 fetch(const float4* __restrict__ f, int w, int h, int c, int r) {
  if (c < 0 || r < 0 || c >= w || r >= h) return make_float4(0.f, 0.f, 0.f, 0.f);
  return f[size_t(r) * w + c];
}

__device__ __forceinline__ void acc4(float4& a, float w, float4 v) {
  a.x += w * v.x;
  a.y += w * v.y;
  a.z += w * v.z;
  a.w += w * v.w;
}

__global__ __launch_bounds__(256) void k_multisample(const float4* __restrict__ src, int sw, int sh,
                                                     float4* __restrict__ dst, int dw, int dh) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= dw || y >= dh) return;
  const float u = (float(x) + 0.5f) / float(dw) * float(sw) - 0.5f;
  const float v = (float(y) + 0.5f) / float(dh) * float(sh) - 0.5f;
  int x0, x1, y0, y1;
  float fx, fy;
  lin_axis(u, sw, x0, x1, fx);
  lin_axis(v, sh, y0, y1, fy);
  const float4 a = src[size_t(y0) * sw + x0], b = src[size_t(y0) * sw + x1];
  const float4 c = src[size_t(y1) * sw + x0], d = src[size_t(y1) * sw + x1];
  float4 o;
  o.x = (a.x + fx * (b.x - a.x)) + fy * ((c.x + fx * (d.x - c.x)) - (a.x + fx * (b.x - a.x)));
  o.y = (a.y + fx * (b.y - a.y)) + fy * ((c.y + fx * (d.y - c.y)) - (a.y + fx * (b.y - a.y)));
  o.z = (a.z + fx * (b.z - a.z)) + fy * ((c.z + fx * (d.z - c.z)) - (a.z + fx * (b.z - a.z)));
  o.w = (a.w + fx * (b.w - a.w)) + fy * ((c.w + fx * (d.w - c.w)) - (a.w + fx * (b.w - a.w)));
  dst[size_t(y) * dw + x] = o;
}

// downscaling_filter.comp:17-57
__global__ __launch_bounds__(256) void k_downscale(const float4* __restrict__ src, int sw, int sh,
                                                   float4* __restrict__ dst, int dw, int dh, int kern) {
  const int jc = blockIdx.x * 16 + (threadIdx.x & 15), jr = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (jc >= dw || jr >= dh) return;
  const float s_r = float(dh) / float(sh), s_c = float(dw) / float(sw);
  const float kr = 0.5f * kernel_support(kern);
  const float x_r = (float(jr) + 0.5f) / float(dh), x_c = (float(jc) + 0.5f) / float(dw);
  const int il_r = int(ceilf((x_r - kr / float(dh)) * float(sh) - 0.5f));
  const int ir_r = int(floorf((x_r + kr / float(dh)) * float(sh) - 0.5f));
  const int il_c = int(ceilf((x_c - kr / float(dw)) * float(sw) - 0.5f));
  const int ir_c = int(floorf((x_c + kr / float(dw)) * float(sw) - 0.5f));
  float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int i_r = il_r; i_r <= ir_r; ++i_r) {
    const float wr = kernel_weight(kern, (x_r - (float(i_r) + 0.5f) / float(sh)) * float(dh));
    for (int i_c = il_c; i_c <= ir_c; ++i_c)
      acc4(f, wr * kernel_weight(kern, (x_c - (float(i_c) + 0.5f) / float(sw)) * float(dw)),
           fetch(src, sw, sh, i_c, i_r));
  }
  const float s = s_r * s_c;
  dst[size_t(jr) * dw + jc] = make_float4(f.x * s, f.y * s, f.z * s, f.w * s);
}

// upscaling_filter.comp:17-55
__global__ __launch_bounds__(256) void k_upscale(const float4* __restrict__ src, int sw, int sh,
                                                 float4* __restrict__ dst, int dw, int dh, int kern) {
  const int jc = blockIdx.x * 16 + (threadIdx.x & 15), jr = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (jc >= dw || jr >= dh) return;
  const float kr = 0.5f * kernel_support(kern);
  const float xi_r = (float(jr) + 0.5f) / float(dh) * float(sh) - 0.5f;
  const float xi_c = (float(jc) + 0.5f) / float(dw) * float(sw) - 0.5f;
  const int il_r = int(ceilf(xi_r - kr)), ir_r = int(floorf(xi_r + kr));
  const int il_c = int(ceilf(xi_c - kr)), ir_c = int(floorf(xi_c + kr));
  float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int i_r = il_r; i_r <= ir_r; ++i_r) {
    const float wr = kernel_weight(kern, xi_r - float(i_r));
    for (int i_c = il_c; i_c <= ir_c; ++i_c)
      acc4(f, wr * kernel_weight(kern, xi_c - float(i_c)), fetch(src, sw, sh, i_c, i_r));
  }
  dst[size_t(jr) * dw + jc] = f;
}

struct DigitalL {
  int m;
  float L[9];
};

// LU factors of the digital filter: L_0 = 1/(a+1), L_i = 1/(a - L_{i-1}).
static DigitalL digital_factors(int kern) {
  DigitalL d{};
  const double a = kern == CVR_FILTER_CARDINAL_BSPLINE3 ? 4.0 : 3.25;
  d.m = kern == CVR_FILTER_CARDINAL_BSPLINE3 ? 8 : 9;
  double l = 1.0 / (a + 1.0);
  for (int i = 0; i < d.m; ++i) {
    d.L[i] = float(l);
    l = 1.0 / (a - l);
  }
  return d;
}

// linear_solve of one sequence of n float4 with stride `step` (filters/utils.hpp:17-44);
// sequences not longer than m are first grown by reflection, as there.
__device__ void solve_seq(float4* __restrict__ f, int n, size_t step, const DigitalL& D) {
  const int m = D.m;
  const float Li = D.L[m - 1], v_inv = Li / (1.f + Li);
  auto axpy = [](float4 a, float s, float4 b) {  // a - s*b
    return make_float4(a.x - s * b.x, a.y - s * b.y, a.z - s * b.z, a.w - s * b.w);
  };
  auto back = [](float s, float4 a, float4 b) {  // s*(a - b)
    return make_float4(s * (a.x - b.x), s * (a.y - b.y), s * (a.z - b.z), s * (a.w - b.w));
  };
  if (n > m) {
    for (int i = 1; i < m; ++i) f[i * step] = axpy(f[i * step], D.L[i - 1], f[(i - 1) * step]);
    for (int i = m; i < n; ++i) f[i * step] = axpy(f[i * step], Li, f[(i - 1) * step]);
    const float4 t = f[(n - 1) * step];
    f[(n - 1) * step] = make_float4(t.x * v_inv, t.y * v_inv, t.z * v_inv, t.w * v_inv);
    for (int i = n - 2; i >= m - 1; --i) f[i * step] = back(Li, f[i * step], f[(i + 1) * step]);
    for (int i = m - 2; i >= 0; --i) f[i * step] = back(D.L[i], f[i * step], f[(i + 1) * step]);
    return;
  }
  float4 g[32];  // n <= 9: reflections until longer than m (at most 2n <= 18... 4n)
  int nn = n;
  for (int i = 0; i < n; ++i) g[i] = f[i * step];
  while (nn <= m) {
    for (int i = 0; i < nn; ++i) g[nn + i] = g[nn - 1 - i];
    nn *= 2;
  }
  for (int i = 1; i < m; ++i) g[i] = axpy(g[i], D.L[i - 1], g[i - 1]);
  for (int i = m; i < nn; ++i) g[i] = axpy(g[i], Li, g[i - 1]);
  g[nn - 1] = make_float4(g[nn - 1].x * v_inv, g[nn - 1].y * v_inv, g[nn - 1].z * v_inv, g[nn - 1].w * v_inv);
  for (int i = nn - 2; i >= m - 1; --i) g[i] = back(Li, g[i], g[i + 1]);
  for (int i = m - 2; i >= 0; --i) g[i] = back(D.L[i], g[i], g[i + 1]);
  for (int i = 0; i < n; ++i) f[i * step] = g[i];
}

// One lane per row (dir 0: along x) or per column (dir 1: along y).
__global__ __launch_bounds__(256) void k_digital(float4* __restrict__ f, int w, int h, int dir, DigitalL D) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (dir == 0) {
    if (i < h) solve_seq(f + size_t(i) * w, w, 1, D);
  } else {
    if (i < w) solve_seq(f + i, h, size_t(w), D);
  }
}

static void digital_filter(float4* f, int w, int h, int kern, hipStream_t st) {
  const DigitalL D = digital_factors(kern);
  hipLaunchKernelGGL(k_digital, dim3(unsigned((h + 255) / 256)), dim3(256), 0, st, f, w, h, 0, D);
  hipLaunchKernelGGL(k_digital, dim3(unsigned((w + 255) / 256)), dim3(256), 0, st, f, w, h, 1, D);
}

}  // namespace cvr

using namespace cvr;

extern "C" int cvr_multiscale_size(int screen_w, int screen_h, int mw, int mh, int* frame_w, int* frame_h) {
  if (!frame_w || !frame_h || screen_w < 1 || screen_h < 1 || mw == 0 || mh == 0) {
    set_error("cvr_multiscale_size: screen size >= 1 and nonzero multipliers required");
    return CVR_ERR_INVALID;
  }
  // UpdateScreenResolutionMultiScaling (renderoutputframe.cpp:96-112): m > 0 multiplies,
  // m < 0 divides (integer division)
  const long long w = mw > 0 ? (long long)screen_w * mw : screen_w / (-(long long)mw);
  const long long h = mh > 0 ? (long long)screen_h * mh : screen_h / (-(long long)mh);
  if (w < 1 || h < 1 || w > (1 << 16) || h > (1 << 16)) {
    set_error("cvr_multiscale_size: frame size out of range");
    return CVR_ERR_INVALID;
  }
  *frame_w = int(w);
  *frame_h = int(h);
  return CVR_OK;
}

extern "C" int cvr_resample_frame(int device, const float* src, int sw, int sh, float* dst, int dw, int dh, int mode,
                                  int kernel, float* work, void* hip_stream) {
  if (!src || !dst || sw < 1 || sh < 1 || dw < 1 || dh < 1 || sw > (1 << 16) || sh > (1 << 16) || dw > (1 << 16) ||
      dh > (1 << 16)) {
    set_error("cvr_resample_frame: device buffers and sizes in [1, 65536] required");
    return CVR_ERR_INVALID;
  }
  if (mode < CVR_SCALE_MULTISAMPLE || mode > CVR_SCALE_UPSCALE || kernel < CVR_FILTER_BOX ||
      kernel > CVR_FILTER_CARDINAL_OMOMS3) {
    set_error("cvr_resample_frame: unknown mode or kernel");
    return CVR_ERR_INVALID;
  }
  const bool cardinal = kernel == CVR_FILTER_CARDINAL_BSPLINE3 || kernel == CVR_FILTER_CARDINAL_OMOMS3;
  if (mode == CVR_SCALE_UPSCALE && cardinal && !work) {
    set_error("cvr_resample_frame: upscaling with a cardinal kernel needs a work buffer of sw*sh*4 floats");
    return CVR_ERR_INVALID;
  }
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) prev = -1;
  if (prev != device) CVR_HIP_CHECK(hipSetDevice(device));
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  const dim3 grid(unsigned((dw + 15) / 16), unsigned((dh + 15) / 16));
  const float4* s4 = reinterpret_cast<const float4*>(src);
  float4* d4 = reinterpret_cast<float4*>(dst);
  if (mode == CVR_SCALE_MULTISAMPLE) {
    hipLaunchKernelGGL(k_multisample, grid, dim3(256), 0, st, s4, sw, sh, d4, dw, dh);
  } else if (mode == CVR_SCALE_DOWNSCALE) {
    hipLaunchKernelGGL(k_downscale, grid, dim3(256), 0, st, s4, sw, sh, d4, dw, dh, kernel);
    if (cardinal) digital_filter(d4, dw, dh, kernel, st);
  } else {
    if (cardinal) {
      float4* w4 = reinterpret_cast<float4*>(work);
      if (work != src) CVR_HIP_CHECK(hipMemcpyAsync(w4, s4, size_t(sw) * sh * sizeof(float4), hipMemcpyDeviceToDevice, st));
      digital_filter(w4, sw, sh, kernel, st);
      s4 = w4;
    }
    hipLaunchKernelGGL(k_upscale, grid, dim3(256), 0, st, s4, sw, sh, d4, dw, dh, kernel);
  }
  const hipError_t e = hipGetLastError();
  if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
  if (e != hipSuccess) {
    set_error(std::string("cvr_resample_frame: ") + hipGetErrorString(e));
    return CVR_ERR_HIP;
  }
  return CVR_OK;
}
