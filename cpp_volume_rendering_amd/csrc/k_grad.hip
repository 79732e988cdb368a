// Volume gradient for Blinn-Phong shading (gfx950).
//
// DataManager::GenerateStructuredGradientTexture (datamanager.cpp:332-352) offers:
//  - Sobel-Feldman (utils.cpp:287-350; the compute-shader form
//    _gradient_shading/sobelfeldman_generator.comp is the same sum): per axis the
//    3x3 weights 4 / 2^(|v1|+|v2|) (4, 2, 1) times (sample on the minus side - sample
//    on the plus side), unnormalised, samples outside the volume 0;
//  - central finite differences with GenerateGradientTexture's defaults
//    (utils.cpp:146-240: step 1, no filter, normalised): normalize(s(+1) - s(-1)) per
//    axis, 0 outside the volume, and the zero vector where the difference vanishes
//    (the reference's NaN test).
// Samples are v/255 or v/65535 (GetNormalizedSample, structuredgridvolume.cpp:121-151).
// One lane per voxel; the grid is written as float4 in the volume's brick order so
// that the 8 taps of a lookup share cache lines.
#include "cvr_ctx.hpp"

namespace cvr {

template <typename T>
__device__ __forceinline__ float norm_sample(const T* __restrict__ v, int nx, int ny, int nz, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return 0.f;
  return float(v[size_t(x) + size_t(nx) * (size_t(y) + size_t(ny) * size_t(z))]) *
         (sizeof(T) == 1 ? (1.0f / 255.0f) : (1.0f / 65535.0f));
}

template <typename T>
__global__ __launch_bounds__(256) void k_gradient(const T* __restrict__ vox, int nx, int ny, int nz, int type, int nbx,
                                                  int nby, float4* __restrict__ out) {
  const size_t n = size_t(nx) * ny * nz;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = int(i % size_t(nx)), y = int((i / size_t(nx)) % size_t(ny)), z = int(i / (size_t(nx) * ny));
  auto s = [&](int a, int b, int c) { return norm_sample<T>(vox, nx, ny, nz, a, b, c); };
  float gx = 0.f, gy = 0.f, gz = 0.f;
  if (type == CVR_GRADIENT_SOBEL) {
    for (int v1 = -1; v1 <= 1; ++v1)
      for (int v2 = -1; v2 <= 1; ++v2) {
        const float w = 4.0f / float(1 << (abs(v1) + abs(v2)));
        gz += s(x + v1, y + v2, z - 1) * w + s(x + v1, y + v2, z + 1) * -w;
        gy += s(x + v1, y - 1, z + v2) * w + s(x + v1, y + 1, z + v2) * -w;
        gx += s(x - 1, y + v2, z + v1) * w + s(x + 1, y + v2, z + v1) * -w;
      }
  } else {
    gx = s(x + 1, y, z) - s(x - 1, y, z);
    gy = s(x, y + 1, z) - s(x, y - 1, z);
    gz = s(x, y, z + 1) - s(x, y, z - 1);
    const float l2 = gx * gx + gy * gy + gz * gz;
    if (l2 > 0.f) {
      const float inv = 1.0f / sqrtf(l2);
      gx *= inv;
      gy *= inv;
      gz *= inv;
    }
  }
  out[brick_index(x, y, z, nbx, nby)] = make_float4(gx, gy, gz, 0.f);
}

int build_gradient(cvr_ctx* c, int type, hipStream_t st) {
  const uint64_t key = (c->vol_version << 8) | uint64_t(type);
  if (c->d_grad && c->grad_built_for == key) return 0;
  if (!c->d_grad) {
    const size_t cells = brick_cells(c->nx, c->ny, c->nz);
    if (hipMalloc(&c->d_grad, cells * sizeof(float4)) != hipSuccess) {
      c->d_grad = nullptr;
      set_error("gradient: out of device memory");
      return CVR_ERR_NOMEM;
    }
    // bricks past the volume's edge are never read; zero them once
    CVR_HIP_CHECK(hipMemsetAsync(c->d_grad, 0, cells * sizeof(float4), st));
  }
  const size_t n = size_t(c->nx) * c->ny * c->nz;
  const dim3 grid(unsigned((n + 255) / 256));
  if (c->dtype == CVR_U8)
    hipLaunchKernelGGL(k_gradient<uint8_t>, grid, dim3(256), 0, st, static_cast<const uint8_t*>(c->d_vol), c->nx, c->ny,
                       c->nz, type, bricks(c->nx), bricks(c->ny), c->d_grad);
  else
    hipLaunchKernelGGL(k_gradient<uint16_t>, grid, dim3(256), 0, st, static_cast<const uint16_t*>(c->d_vol), c->nx,
                       c->ny, c->nz, type, bricks(c->nx), bricks(c->ny), c->d_grad);
  CVR_HIP_CHECK(hipGetLastError());
  c->grad_built_for = key;
  return 0;
}

}  // namespace cvr
