// Gaussian extinction mip pyramid on the GPU, gfx950.
//
// Follows rc1pdosct/extcoefvolumegenerator.cpp:230-408 and its shaders
// glslextgen/gen_extcoefvol_anysize.comp (level 0: 7^3 Gaussian taps of the TF
// opacity of the trilinear volume sample, sigma S0, zero outside the volume),
// gen_extcoefvol_anysize_mmlevel.comp (level i: 7^3 taps of level i-1 with sigma
// S0*2^i) and backtotau.comp (extinction = -log(1 - opacity) per level).
// Levels are stored in one fp32 buffer, level 0 first, x fastest.  Opacity levels
// live in a scratch buffer only while the chain is built.
//
// Direct build: one lane per output voxel, 8x8x4 voxel bricks per 256-lane workgroup;
// the 343 weights are the same for every voxel and sit in LDS.  Separable build (used
// when the voxel spacing and sigma are commensurate, as in every default configuration):
// see k_pyr_lattice / k_pyr_sep below.
#include "cvr_ctx.hpp"
#include <array>
#include <algorithm>
#include <cstring>
#include <cmath>
#include <vector>

namespace cvr {

struct PyrBuild {
  int bx, by, bz;  // output level dims
  int px, py, pz;  // previous level dims (levels >= 1)
  float G[3], vs[3];
  float sigma;
};

__device__ __forceinline__ bool outside01(float x, float y, float z) {
  return x < 0.f || y < 0.f || z < 0.f || x > 1.f || y > 1.f || z > 1.f;
}

template <typename T>
__global__ __launch_bounds__(256) void k_pyr_level0(PyrBuild b, VolGeom g, const float4* __restrict__ tf_rgba,
                                                     int ntf, float* __restrict__ out_op) {
  __shared__ float s_w[343];
  __shared__ float s_wsum;
  if (threadIdx.x == 0) s_wsum = 0.f;
  __syncthreads();
  for (int i = threadIdx.x; i < 343; i += blockDim.x) {
    const int kx = i % 7 - 3, ky = (i / 7) % 7 - 3, kz = i / 49 - 3;
    const float fx = float(kx) * b.sigma, fy = float(ky) * b.sigma, fz = float(kz) * b.sigma;
    const float w = (b.sigma * b.sigma * b.sigma) * expf(-(fx * fx + fy * fy + fz * fz) / (2.0f * b.sigma * b.sigma));
    s_w[i] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < 343; ++i) s += s_w[i];
    s_wsum = s;
  }
  __syncthreads();
  const int x = blockIdx.x * 8 + (threadIdx.x & 7);
  const int y = blockIdx.y * 8 + ((threadIdx.x >> 3) & 7);
  const int z = blockIdx.z * 4 + (threadIdx.x >> 6);
  if (x >= b.bx || y >= b.by || z >= b.bz) return;
  const float gx = (float(x) + 0.5f) * b.vs[0], gy = (float(y) + 0.5f) * b.vs[1], gz = (float(z) + 0.5f) * b.vs[2];
  float acc = 0.f;
  for (int i = 0; i < 343; ++i) {
    const int kx = i % 7 - 3, ky = (i / 7) % 7 - 3, kz = i / 49 - 3;
    const float px = gx + float(kx) * b.sigma, py = gy + float(ky) * b.sigma, pz = gz + float(kz) * b.sigma;
    const float tx = px / b.G[0], ty = py / b.G[1], tz = pz / b.G[2];
    if (outside01(tx, ty, tz)) continue;
    const float d = sample_volume_oct<T>(g, mk3(px, py, pz));
    acc += s_w[i] * sample_tf(tf_rgba, ntf, d).w;
  }
  out_op[size_t(x) + size_t(b.bx) * (y + size_t(b.by) * z)] = acc / s_wsum;
}

// trilinear (GL_LINEAR, clamp) over a float grid at normalised coordinate (u,v,w)
__device__ __forceinline__ float tex3_f(const float* __restrict__ a, int nx, int ny, int nz, float u, float v,
                                        float w) {
  int x0, x1, y0, y1, z0, z1;
  float fx, fy, fz;
  lin_axis(u * float(nx) - 0.5f, nx, x0, x1, fx);
  lin_axis(v * float(ny) - 0.5f, ny, y0, y1, fy);
  lin_axis(w * float(nz) - 0.5f, nz, z0, z1, fz);
  const size_t sy = size_t(nx), sz = size_t(nx) * ny;
  const size_t r00 = y0 * sy + z0 * sz, r10 = y1 * sy + z0 * sz, r01 = y0 * sy + z1 * sz, r11 = y1 * sy + z1 * sz;
  const float c00 = a[r00 + x0] + fx * (a[r00 + x1] - a[r00 + x0]);
  const float c10 = a[r10 + x0] + fx * (a[r10 + x1] - a[r10 + x0]);
  const float c01 = a[r01 + x0] + fx * (a[r01 + x1] - a[r01 + x0]);
  const float c11 = a[r11 + x0] + fx * (a[r11 + x1] - a[r11 + x0]);
  const float c0 = c00 + fy * (c10 - c00), c1 = c01 + fy * (c11 - c01);
  return c0 + fz * (c1 - c0);
}

__global__ __launch_bounds__(256) void k_pyr_next(PyrBuild b, const float* __restrict__ prev_op,
                                                  float* __restrict__ out_op) {
  __shared__ float s_w[343];
  __shared__ float s_wsum;
  for (int i = threadIdx.x; i < 343; i += blockDim.x) {
    const int kx = i % 7 - 3, ky = (i / 7) % 7 - 3, kz = i / 49 - 3;
    const float fx = float(kx) * b.sigma, fy = float(ky) * b.sigma, fz = float(kz) * b.sigma;
    s_w[i] = (b.sigma * b.sigma * b.sigma) * expf(-(fx * fx + fy * fy + fz * fz) / (2.0f * b.sigma * b.sigma));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < 343; ++i) s += s_w[i];
    s_wsum = s;
  }
  __syncthreads();
  const int x = blockIdx.x * 8 + (threadIdx.x & 7);
  const int y = blockIdx.y * 8 + ((threadIdx.x >> 3) & 7);
  const int z = blockIdx.z * 4 + (threadIdx.x >> 6);
  if (x >= b.bx || y >= b.by || z >= b.bz) return;
  const float gx = (float(x) + 0.5f) * b.vs[0], gy = (float(y) + 0.5f) * b.vs[1], gz = (float(z) + 0.5f) * b.vs[2];
  float acc = 0.f;
  for (int i = 0; i < 343; ++i) {
    const int kx = i % 7 - 3, ky = (i / 7) % 7 - 3, kz = i / 49 - 3;
    const float tx = (gx + float(kx) * b.sigma) / b.G[0];
    const float ty = (gy + float(ky) * b.sigma) / b.G[1];
    const float tz = (gz + float(kz) * b.sigma) / b.G[2];
    if (outside01(tx, ty, tz)) continue;
    acc += s_w[i] * tex3_f(prev_op, b.px, b.py, b.pz, tx, ty, tz);
  }
  out_op[size_t(x) + size_t(b.bx) * (y + size_t(b.by) * z)] = acc / s_wsum;
}

// ---------------------------------------------------------------------------------
// Separable build.  The 7^3 taps of output voxel x along axis a sit at (x + 1/2) vs_a +
// i sigma (i = -3..3).  When vs_a = m_a delta_a and sigma = q_a delta_a for small
// integers m_a, q_a, every tap of every output voxel lies on the lattice (1/2) vs_a +
// (n - 3 q_a) delta_a, n = m_a x + q_a (i + 3), and the weights factor as
// sigma^3 g_i g_j g_k (g_i = exp(-i^2 / 2)); the zero outside [0, G] is a per-axis test.
// So a level is: its function (TF opacity of the trilinear volume sample for level 0,
// the trilinear previous level above) evaluated once per lattice point, then three 7-tap
// passes along x, y and z with output stride m_a and tap stride q_a, normalised by
// (sum g)^3 -- 3 x 7 taps per voxel instead of 343, the same samples and weights
// (extcoefvolumegenerator.cpp:230-367).  vs / sigma is the same at every level
// (both double per level), so the whole chain takes this path or none of it.
struct SepAxis {
  int m, q, len;  // output stride, tap stride, lattice points (m (b - 1) + 6 q + 1)
  float half_vs, delta;
};

template <typename T, bool kLevel0>
__global__ __launch_bounds__(256) void k_pyr_lattice(SepAxis ax, SepAxis ay, SepAxis az, float3 G, VolGeom g,
                                                     const float4* __restrict__ tf_rgba, int ntf,
                                                     const float* __restrict__ prev_op, int px, int py, int pz,
                                                     float* __restrict__ lat) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  const size_t n = size_t(ax.len) * ay.len * az.len;
  if (i >= n) return;
  const int nx = int(i % ax.len), ny = int((i / ax.len) % ay.len), nz = int(i / (size_t(ax.len) * ay.len));
  const float p0 = ax.half_vs + float(nx - 3 * ax.q) * ax.delta;
  const float p1 = ay.half_vs + float(ny - 3 * ay.q) * ay.delta;
  const float p2 = az.half_vs + float(nz - 3 * az.q) * az.delta;
  const float tx = p0 / G.x, ty = p1 / G.y, tz = p2 / G.z;
  float v = 0.f;
  if (!outside01(tx, ty, tz)) {
    if (kLevel0) v = sample_tf(tf_rgba, ntf, sample_volume_oct<T>(g, mk3(p0, p1, p2))).w;
    else v = tex3_f(prev_op, px, py, pz, tx, ty, tz);
  }
  lat[i] = v;
}

// One separable pass along `axis` of an (in0 x in1 x in2) grid (x fastest): output
// index o along the axis takes the 7 inputs m o + q t (t = 0..6) with weights g[t];
// the other two axes pass through.  `scale` multiplies the sum (the normalisation on
// the last pass).
struct SepW {
  float g[7];
};
__global__ __launch_bounds__(256) void k_pyr_sep(const float* __restrict__ in, float* __restrict__ out, int in0, int in1,
                                                 int in2, int axis, int m, int q, int outlen, SepW w, float scale) {
  const int o0 = axis == 0 ? outlen : in0, o1 = axis == 1 ? outlen : in1, o2 = axis == 2 ? outlen : in2;
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= size_t(o0) * o1 * o2) return;
  const int x = int(i % o0), y = int((i / o0) % o1), z = int(i / (size_t(o0) * o1));
  const size_t s1 = size_t(in0), s2 = size_t(in0) * in1;
  size_t base, step;
  if (axis == 0) {
    base = size_t(m) * x + y * s1 + z * s2;
    step = size_t(q);
  } else if (axis == 1) {
    base = x + size_t(m) * y * s1 + z * s2;
    step = size_t(q) * s1;
  } else {
    base = x + y * s1 + size_t(m) * z * s2;
    step = size_t(q) * s2;
  }
  float acc = 0.f;
#pragma unroll
  for (int t = 0; t < 7; ++t) acc += w.g[t] * in[base + t * step];
  out[i] = acc * scale;
}

// Small integers m, q with vs = m delta and sigma = q delta (to 1e-6 relative), or false.
static bool sep_axis(double vs, double sigma, int b, SepAxis& a) {
  for (int q = 1; q <= 8; ++q)
    for (int m = 1; m <= 16; ++m)
      if (std::fabs(double(m) * sigma - double(q) * vs) <= 1e-6 * std::max(vs, sigma)) {
        a.m = m;
        a.q = q;
        a.len = m * (b - 1) + 6 * q + 1;
        a.half_vs = float(0.5 * vs);
        a.delta = float(sigma / q);
        return true;
      }
  return false;
}

// backtotau.comp:11-35
__global__ __launch_bounds__(256) void k_pyr_tau(const float* __restrict__ op, float* __restrict__ tau, size_t n) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) tau[i] = -1.0f * logf(1.0f - op[i]);
}

static int mip_count(int bx, int by, int bz) {
  int m = std::max(bx, std::max(by, bz)), c = 1;
  while (m > 1) { m >>= 1; ++c; }
  return c;
}

PyrKey pyr_key_for(const cvr_ctx* c, int base_x, int base_y, int base_z, float sigma0, int nlev, int fmt) {
  PyrKey key{};
  key.vol_v = c->vol_version;
  key.tf_v = c->tf_version;
  key.base[0] = base_x; key.base[1] = base_y; key.base[2] = base_z;
  std::memcpy(&key.sigma_bits, &sigma0, 4);
  key.nlev = nlev;
  key.fmt = fmt;
  return key;
}

int build_pyramid(cvr_ctx* c, int base_x, int base_y, int base_z, float sigma0, int nlev, int fmt, hipStream_t st) {
  if (base_x < 1 || base_y < 1 || base_z < 1) { set_error("pyramid: base resolution must be >= 1"); return -1; }
  if (!(sigma0 > 0)) { set_error("pyramid: sigma0 must be > 0"); return -1; }
  const PyrKey key = pyr_key_for(c, base_x, base_y, base_z, sigma0, nlev, fmt);
  if (fmt < 0 || fmt > 2) { set_error("pyramid: density_precision must be 0 (fp32), 1 (fp16) or 2 (bf16)"); return -1; }
  if (c->d_pyr && c->pyr_valid && c->pyr_key == key) return 0;
  c->pyr_valid = false;
  const int full = mip_count(base_x, base_y, base_z);
  const int L = nlev > 0 ? std::min(nlev, full) : full;
  if (L > kMaxLevels) { set_error("pyramid: too many levels"); return -4; }
  PyrGeom pg{};
  pg.levels = L;
  pg.fmt = fmt;
  size_t total = 0, btotal = 0;
  for (int i = 0; i < L; ++i) {
    pg.dims[i][0] = std::max(1, base_x >> i);
    pg.dims[i][1] = std::max(1, base_y >> i);
    pg.dims[i][2] = std::max(1, base_z >> i);
    pg.off[i] = total;
    pg.boff[i] = btotal;
    for (int k = 0; k < 3; ++k) {
      pg.ts[i][k] = float(double(pg.dims[i][k]) / double(geom_of(c).G[k]));
      pg.dmax[i][k] = float(pg.dims[i][k] - 1);
    }
    pg.nb[i][0] = bricks(pg.dims[i][0]);
    pg.nb[i][1] = bricks(pg.dims[i][1]);
    btotal += brick_cells(pg.dims[i][0], pg.dims[i][1], pg.dims[i][2]);
    if (brick_cells(pg.dims[i][0], pg.dims[i][1], pg.dims[i][2]) * (fmt == 0 ? 32u : 16u) > 0xFFFFFFFFull) pg.wide = 1;
    total += size_t(pg.dims[i][0]) * pg.dims[i][1] * pg.dims[i][2];
  }
  VolGeom g = geom_of(c);
  for (int k = 0; k < 3; ++k) pg.G[k] = g.G[k];
  pg.one = 1.0f;
  if (c->d_pyr) CVR_HIP_CHECK(hipFree(c->d_pyr));
  c->d_pyr = nullptr;
  float* d_op = nullptr;
  if (hipMalloc(&c->d_pyr, total * sizeof(float)) != hipSuccess || hipMalloc(&d_op, total * sizeof(float)) != hipSuccess) {
    if (c->d_pyr) (void)hipFree(c->d_pyr);
    c->d_pyr = nullptr;
    set_error("pyramid: out of device memory");
    return -5;
  }
  // separable path per level: voxel spacing and sigma commensurate on every axis (true at
  // every level of the default configurations; a level whose size was rounded down from
  // an odd one may not be, and takes the direct build)
  std::vector<std::array<SepAxis, 3>> sax(L);
  std::vector<char> sep(L, 0);
  size_t tmp_floats = 0;
  for (int i = 0; i < L; ++i) {
    bool ok = true;
    const double sig = double(sigma0) * std::pow(2.0, double(i));
    for (int k = 0; k < 3 && ok; ++k) ok = sep_axis(double(g.G[k]) / double(pg.dims[i][k]), sig, pg.dims[i][k], sax[i][k]);
    sep[i] = ok;
    if (ok)
      tmp_floats = std::max(tmp_floats, size_t(sax[i][0].len) * sax[i][1].len * sax[i][2].len +
                                            size_t(pg.dims[i][0]) * sax[i][1].len * sax[i][2].len);
  }
  float* d_tmp = nullptr;
  if (tmp_floats && hipMalloc(&d_tmp, tmp_floats * sizeof(float)) != hipSuccess) {
    (void)hipGetLastError();
    d_tmp = nullptr;
    std::fill(sep.begin(), sep.end(), 0);  // no room for the temporaries: the direct build needs none
  }
  SepW gw;
  double gsum = 0;
  for (int t = 0; t < 7; ++t) {
    gw.g[t] = float(std::exp(-0.5 * double((t - 3) * (t - 3))));
    gsum += double(gw.g[t]);
  }
  const float norm = float(1.0 / (gsum * gsum * gsum));
  const float3 G3 = make_float3(g.G[0], g.G[1], g.G[2]);
  for (int i = 0; i < L; ++i) {
    if (!sep[i]) {
      // direct build of this level (343 taps per voxel)
      PyrBuild b{};
      b.bx = pg.dims[i][0]; b.by = pg.dims[i][1]; b.bz = pg.dims[i][2];
      for (int k = 0; k < 3; ++k) {
        b.G[k] = g.G[k];
        b.vs[k] = g.G[k] / float(pg.dims[i][k]);
      }
      b.sigma = sigma0 * float(std::pow(2.0, double(i)));
      dim3 grid((b.bx + 7) / 8, (b.by + 7) / 8, (b.bz + 3) / 4);
      if (i == 0) {
        if (c->dtype == CVR_U8)
          hipLaunchKernelGGL(k_pyr_level0<uint8_t>, grid, dim3(256), 0, st, b, g, c->d_tf_rgba, c->tf_size, d_op);
        else
          hipLaunchKernelGGL(k_pyr_level0<uint16_t>, grid, dim3(256), 0, st, b, g, c->d_tf_rgba, c->tf_size, d_op);
      } else {
        b.px = pg.dims[i - 1][0]; b.py = pg.dims[i - 1][1]; b.pz = pg.dims[i - 1][2];
        hipLaunchKernelGGL(k_pyr_next, grid, dim3(256), 0, st, b, d_op + pg.off[i - 1], d_op + pg.off[i]);
      }
      CVR_HIP_CHECK(hipGetLastError());
      continue;
    }
    const std::array<SepAxis, 3>& a = sax[i];
    const size_t nlat = size_t(a[0].len) * a[1].len * a[2].len;
    float* lat = d_tmp;
    float* p1 = d_tmp + nlat;
    const unsigned nb = unsigned((nlat + 255) / 256);
    if (i == 0) {
      if (c->dtype == CVR_U8)
        hipLaunchKernelGGL((k_pyr_lattice<uint8_t, true>), dim3(nb), dim3(256), 0, st, a[0], a[1], a[2], G3, g,
                           c->d_tf_rgba, c->tf_size, nullptr, 0, 0, 0, lat);
      else
        hipLaunchKernelGGL((k_pyr_lattice<uint16_t, true>), dim3(nb), dim3(256), 0, st, a[0], a[1], a[2], G3, g,
                           c->d_tf_rgba, c->tf_size, nullptr, 0, 0, 0, lat);
    } else {
      hipLaunchKernelGGL((k_pyr_lattice<uint8_t, false>), dim3(nb), dim3(256), 0, st, a[0], a[1], a[2], G3, g,
                         nullptr, 0, d_op + pg.off[i - 1], pg.dims[i - 1][0], pg.dims[i - 1][1], pg.dims[i - 1][2], lat);
    }
    CVR_HIP_CHECK(hipGetLastError());
    const int bx = pg.dims[i][0], by = pg.dims[i][1], bz = pg.dims[i][2];
    const size_t n1 = size_t(bx) * a[1].len * a[2].len, n2 = size_t(bx) * by * a[2].len, n3 = size_t(bx) * by * bz;
    // x: lattice -> p1; y: p1 -> the lattice's space; z: -> the level, normalised
    hipLaunchKernelGGL(k_pyr_sep, dim3(unsigned((n1 + 255) / 256)), dim3(256), 0, st, lat, p1, a[0].len, a[1].len,
                       a[2].len, 0, a[0].m, a[0].q, bx, gw, 1.0f);
    hipLaunchKernelGGL(k_pyr_sep, dim3(unsigned((n2 + 255) / 256)), dim3(256), 0, st, p1, lat, bx, a[1].len, a[2].len,
                       1, a[1].m, a[1].q, by, gw, 1.0f);
    hipLaunchKernelGGL(k_pyr_sep, dim3(unsigned((n3 + 255) / 256)), dim3(256), 0, st, lat, d_op + pg.off[i], bx, by,
                       a[2].len, 2, a[2].m, a[2].q, bz, gw, norm);
    CVR_HIP_CHECK(hipGetLastError());
  }
  if (d_tmp) {
    CVR_HIP_CHECK(hipStreamSynchronize(st));
    CVR_HIP_CHECK(hipFree(d_tmp));
  }
  hipLaunchKernelGGL(k_pyr_tau, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, d_op, c->d_pyr, total);
  CVR_HIP_CHECK(hipGetLastError());
  CVR_HIP_CHECK(hipStreamSynchronize(st));
  CVR_HIP_CHECK(hipFree(d_op));
  c->pyr = pg;
  if (build_pyramid_oct(c, btotal, fmt, st) != CVR_OK) return -5;
  c->pyr_key = key;
  c->pyr_valid = true;
  return 0;
}

}  // namespace cvr
