// Neighbour-block ("quad") copies of the sampled grids.  Cell (x, y, z) of a quad grid
// holds the values at (x, y), (x+1, y), (x, y+1), (x+1, y+1) of plane z, with the +1
// neighbours clamped to the edge, so a GL_LINEAR trilinear read is two vector loads
// (planes z0 and z1) instead of eight scalar gathers.  Used for the raw volume
// (uchar4 / ushort4 per voxel) and for every level of the extinction pyramid (float4).
#include "cvr_ctx.hpp"

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
  dst[r0 + x] = q;
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
  const size_t cells = size_t(c->nx) * c->ny * c->nz;
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

int build_pyramid_quad(cvr_ctx* c, size_t total, hipStream_t st) {
  if (c->d_pyr_quad) CVR_HIP_CHECK(hipFree(c->d_pyr_quad));
  c->d_pyr_quad = nullptr;
  if (hipMalloc(&c->d_pyr_quad, total * sizeof(float4)) != hipSuccess) {
    c->d_pyr_quad = nullptr;
    set_error("pyramid: out of device memory (quad layout)");
    return CVR_ERR_NOMEM;
  }
  const PyrGeom& pg = c->pyr;
  for (int l = 0; l < pg.levels; ++l) {
    const int rc = make_quad(c->d_pyr + pg.off[l], c->d_pyr_quad + pg.off[l], pg.dims[l][0], pg.dims[l][1],
                             pg.dims[l][2], st);
    if (rc != CVR_OK) return rc;
  }
  CVR_HIP_CHECK(hipStreamSynchronize(st));
  return CVR_OK;
}

}  // namespace cvr
