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
// One lane per output voxel, 8x8x4 voxel bricks per 256-lane workgroup; the 343
// weights are the same for every voxel and sit in LDS.
#include "cvr_ctx.hpp"
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

int build_pyramid(cvr_ctx* c, int base_x, int base_y, int base_z, float sigma0, int nlev, int fmt, hipStream_t st) {
  if (base_x < 1 || base_y < 1 || base_z < 1) { set_error("pyramid: base resolution must be >= 1"); return -1; }
  if (!(sigma0 > 0)) { set_error("pyramid: sigma0 must be > 0"); return -1; }
  PyrKey key{};
  key.vol_v = c->vol_version;
  key.tf_v = c->tf_version;
  key.base[0] = base_x; key.base[1] = base_y; key.base[2] = base_z;
  std::memcpy(&key.sigma_bits, &sigma0, 4);
  key.nlev = nlev;
  key.fmt = fmt;
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
    for (int k = 0; k < 3; ++k) pg.ts[i][k] = float(double(pg.dims[i][k]) / double(geom_of(c).G[k]));
    pg.nb[i][0] = bricks(pg.dims[i][0]);
    pg.nb[i][1] = bricks(pg.dims[i][1]);
    btotal += brick_cells(pg.dims[i][0], pg.dims[i][1], pg.dims[i][2]);
    total += size_t(pg.dims[i][0]) * pg.dims[i][1] * pg.dims[i][2];
  }
  VolGeom g = geom_of(c);
  for (int k = 0; k < 3; ++k) pg.G[k] = g.G[k];
  if (c->d_pyr) CVR_HIP_CHECK(hipFree(c->d_pyr));
  c->d_pyr = nullptr;
  float* d_op = nullptr;
  if (hipMalloc(&c->d_pyr, total * sizeof(float)) != hipSuccess || hipMalloc(&d_op, total * sizeof(float)) != hipSuccess) {
    if (c->d_pyr) (void)hipFree(c->d_pyr);
    c->d_pyr = nullptr;
    set_error("pyramid: out of device memory");
    return -5;
  }
  for (int i = 0; i < L; ++i) {
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
