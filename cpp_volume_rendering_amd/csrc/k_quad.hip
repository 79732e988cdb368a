// Neighbour-block ("quad") copies of the sampled grids.  Cell (x, y, z) of a quad grid
// holds the values at (x, y), (x+1, y), (x, y+1), (x+1, y+1) of plane z, with the +1
// neighbours clamped to the edge, so a GL_LINEAR trilinear read is two vector loads
// (planes z0 and z1) instead of eight scalar gathers.  Used for the raw volume
// (uchar4 / ushort4 per voxel) and for every level of the extinction pyramid (float4,
// or four fp16 / bf16 values in 8 bytes).  The blocks are stored in brick order
// (4 x 4 x 4 cells per brick, brick_index), so that the samples of a wavefront, which
// spread over a small 3D region, touch few cache lines.
#include "cvr_ctx.hpp"
#include <hip/hip_fp16.h>

namespace cvr {

template <typename T, typename Q>
__global__ __launch_bounds__(256) void k_make_quad(const T* __restrict__ src, Q* __restrict__ dst, int nx, int ny,
                                                   int nz) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int z = blockIdx.z;
  if (x >= nx || y >= ny || z >= nz) return;
  const int x1 = min(x + 1, nx - 1), y1 = min(y + 1, ny - 1);
  const size_t r0 = (size_t(z) * ny + y) * nx, r1 = (size_t(z) * ny + y1) * nx;
  Q q;
  q.x = src[r0 + x];
  q.y = src[r0 + x1];
  q.z = src[r1 + x];
  q.w = src[r1 + x1];
  dst[brick_index(x, y, z, bricks(nx), bricks(ny))] = q;
}

template <typename T, typename Q>
static int make_quad(const T* src, Q* dst, int nx, int ny, int nz, hipStream_t st) {
  if (nz > 65535) { set_error("quad layout: depth above 65535"); return CVR_ERR_INVALID; }
  const dim3 grid(unsigned((nx + 63) / 64), unsigned((ny + 3) / 4), unsigned(nz));
  hipLaunchKernelGGL((k_make_quad<T, Q>), grid, dim3(256), 0, st, src, dst, nx, ny, nz);
  CVR_HIP_CHECK(hipGetLastError());
  return CVR_OK;
}

int build_volume_quad(cvr_ctx* c, hipStream_t st) {
  const size_t cells = brick_cells(c->nx, c->ny, c->nz);
  const size_t bytes = cells * (c->dtype == CVR_U8 ? sizeof(uchar4) : sizeof(ushort4));
  if (c->d_vol_quad) CVR_HIP_CHECK(hipFree(c->d_vol_quad));
  c->d_vol_quad = nullptr;
  if (hipMalloc(&c->d_vol_quad, bytes) != hipSuccess) {
    c->d_vol_quad = nullptr;
    set_error("cvr_set_volume: out of device memory (quad layout)");
    return CVR_ERR_NOMEM;
  }
  int rc;
  if (c->dtype == CVR_U8)
    rc = make_quad(static_cast<const uint8_t*>(c->d_vol), static_cast<uchar4*>(c->d_vol_quad), c->nx, c->ny, c->nz, st);
  else
    rc = make_quad(static_cast<const uint16_t*>(c->d_vol), static_cast<ushort4*>(c->d_vol_quad), c->nx, c->ny, c->nz,
                   st);
  if (rc != CVR_OK) return rc;
  CVR_HIP_CHECK(hipStreamSynchronize(st));
  return CVR_OK;
}

// fp32 -> 16-bit storage, round to nearest even: fp16 (fmt 1) or bf16 (fmt 2).
__device__ __forceinline__ unsigned int to16(float x, int fmt) {
  if (fmt == 1) return __half_as_ushort(__float2half_rn(x));
  unsigned int u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}

// Narrows fp32 neighbour blocks to 4 x 16 bit.
__global__ __launch_bounds__(256) void k_pack_quad16(const float4* __restrict__ src, uint2* __restrict__ dst, size_t n,
                                                     int fmt) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 q = src[i];
  dst[i] = make_uint2(to16(q.x, fmt) | (to16(q.y, fmt) << 16), to16(q.z, fmt) | (to16(q.w, fmt) << 16));
}

int build_pyramid_quad(cvr_ctx* c, size_t total, int fmt, hipStream_t st) {
  if (c->d_pyr_quad) CVR_HIP_CHECK(hipFree(c->d_pyr_quad));
  c->d_pyr_quad = nullptr;
  float4* q32 = nullptr;
  if (hipMalloc(&q32, total * sizeof(float4)) != hipSuccess) {
    set_error("pyramid: out of device memory (quad layout)");
    return CVR_ERR_NOMEM;
  }
  const PyrGeom& pg = c->pyr;
  for (int l = 0; l < pg.levels; ++l) {
    const int rc = make_quad(c->d_pyr + pg.off[l], q32 + pg.boff[l], pg.dims[l][0], pg.dims[l][1], pg.dims[l][2], st);
    if (rc != CVR_OK) {
      (void)hipFree(q32);
      return rc;
    }
  }
  if (fmt == 0) {
    CVR_HIP_CHECK(hipStreamSynchronize(st));
    c->d_pyr_quad = q32;
    return CVR_OK;
  }
  void* q16 = nullptr;
  if (hipMalloc(&q16, total * sizeof(uint2)) != hipSuccess) {
    (void)hipFree(q32);
    set_error("pyramid: out of device memory (quad layout)");
    return CVR_ERR_NOMEM;
  }
  hipLaunchKernelGGL(k_pack_quad16, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, q32,
                     static_cast<uint2*>(q16), total, fmt);
  CVR_HIP_CHECK(hipGetLastError());
  CVR_HIP_CHECK(hipStreamSynchronize(st));
  CVR_HIP_CHECK(hipFree(q32));
  c->d_pyr_quad = q16;
  return CVR_OK;
}

}  // namespace cvr
