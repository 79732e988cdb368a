// Corner-block copies of the sampled grids, so that a GL_LINEAR trilinear read is one
// vector load instead of eight scalar gathers: cell (x, y, z) holds all eight values of
// (x..x+1, y..y+1, z..z+1), the +1 neighbours clamped to the edge, plane z first.
//  - Raw volume: 8 x u8 (8 bytes) or 8 x u16 (16 bytes) per cell.
//  - Extinction pyramid, every level: 32 bytes in fp32, 16 in fp16 / bf16.
// A read is one record on one cache line: eight times the raw bytes of the volume (four
// of a 2 x 2 plane block), for half the loads and half the cache lines of plane blocks,
// which is what the marches and the cone-traced lookups are bound by.  Both are stored
// in brick order (4 x 4 x 4 cells per brick, brick_index), so that the samples of a
// wavefront, which spread over a small 3D region, touch few cache lines.
#include "cvr_ctx.hpp"
#include <hip/hip_fp16.h>

namespace cvr {

// Corner blocks of the raw voxels (u8: 8 bytes, u16: 16 bytes per cell).
template <typename T>
__global__ __launch_bounds__(256) void k_make_vol_oct(const T* __restrict__ src, void* __restrict__ dst, int nx, int ny,
                                                      int nz) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int z = blockIdx.z;
  if (x >= nx || y >= ny || z >= nz) return;
  const int x1 = min(x + 1, nx - 1), y1 = min(y + 1, ny - 1), z1 = min(z + 1, nz - 1);
  const size_t sy = size_t(nx), sz = size_t(nx) * ny;
  const size_t r00 = z * sz + y * sy, r01 = z * sz + y1 * sy, r10 = z1 * sz + y * sy, r11 = z1 * sz + y1 * sy;
  const unsigned int v[8] = {src[r00 + x], src[r00 + x1], src[r01 + x], src[r01 + x1],
                             src[r10 + x], src[r10 + x1], src[r11 + x], src[r11 + x1]};
  const size_t i = brick_index(x, y, z, bricks(nx), bricks(ny));
  if (sizeof(T) == 1) {
    static_cast<uint2*>(dst)[i] = make_uint2(v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24),
                                             v[4] | (v[5] << 8) | (v[6] << 16) | (v[7] << 24));
  } else {
    static_cast<uint4*>(dst)[i] = make_uint4(v[0] | (v[1] << 16), v[2] | (v[3] << 16), v[4] | (v[5] << 16), v[6] | (v[7] << 16));
  }
}

template <typename T>
static int make_vol_oct(const T* src, void* dst, int nx, int ny, int nz, hipStream_t st) {
  if (nz > 65535) { set_error("corner blocks: depth above 65535"); return CVR_ERR_INVALID; }
  const dim3 grid(unsigned((nx + 63) / 64), unsigned((ny + 3) / 4), unsigned(nz));
  hipLaunchKernelGGL((k_make_vol_oct<T>), grid, dim3(256), 0, st, src, dst, nx, ny, nz);
  CVR_HIP_CHECK(hipGetLastError());
  return CVR_OK;
}

int build_volume_oct(cvr_ctx* c, hipStream_t st) {
  const size_t cells = brick_cells(c->nx, c->ny, c->nz);
  const size_t bytes = cells * (c->dtype == CVR_U8 ? sizeof(uint2) : sizeof(uint4));
  if (c->d_vol_oct) CVR_HIP_CHECK(hipFree(c->d_vol_oct));
  c->d_vol_oct = nullptr;
  if (hipMalloc(&c->d_vol_oct, bytes) != hipSuccess) {
    c->d_vol_oct = nullptr;
    set_error("cvr_set_volume: out of device memory (corner blocks)");
    return CVR_ERR_NOMEM;
  }
  int rc;
  if (c->dtype == CVR_U8)
    rc = make_vol_oct(static_cast<const uint8_t*>(c->d_vol), c->d_vol_oct, c->nx, c->ny, c->nz, st);
  else
    rc = make_vol_oct(static_cast<const uint16_t*>(c->d_vol), c->d_vol_oct, c->nx, c->ny, c->nz, st);
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

// Corner blocks of one pyramid level: cell (x, y, z) holds the 2 x 2 x 2 values at
// (x..x+1, y..y+1, z..z+1), the +1 neighbours clamped to the edge, plane z first, each
// plane (x, y), (x+1, y), (x, y+1), (x+1, y+1); fp32 as 32 bytes, fp16 / bf16 as 16.
template <int FMT>
__global__ __launch_bounds__(256) void k_make_oct(const float* __restrict__ src, void* __restrict__ dst, int nx, int ny,
                                                  int nz) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int z = blockIdx.z;
  if (x >= nx || y >= ny || z >= nz) return;
  const int x1 = min(x + 1, nx - 1), y1 = min(y + 1, ny - 1), z1 = min(z + 1, nz - 1);
  const size_t sy = size_t(nx), sz = size_t(nx) * ny;
  const size_t r00 = z * sz + y * sy, r01 = z * sz + y1 * sy, r10 = z1 * sz + y * sy, r11 = z1 * sz + y1 * sy;
  const float v[8] = {src[r00 + x], src[r00 + x1], src[r01 + x], src[r01 + x1],
                      src[r10 + x], src[r10 + x1], src[r11 + x], src[r11 + x1]};
  const size_t i = brick_index(x, y, z, bricks(nx), bricks(ny));
  if (FMT == 0) {
    float4* d = static_cast<float4*>(dst) + 2 * i;
    d[0] = make_float4(v[0], v[1], v[2], v[3]);
    d[1] = make_float4(v[4], v[5], v[6], v[7]);
  } else {
    static_cast<uint4*>(dst)[i] = make_uint4(to16(v[0], FMT) | (to16(v[1], FMT) << 16), to16(v[2], FMT) | (to16(v[3], FMT) << 16),
                                             to16(v[4], FMT) | (to16(v[5], FMT) << 16), to16(v[6], FMT) | (to16(v[7], FMT) << 16));
  }
}

int build_pyramid_oct(cvr_ctx* c, size_t total, int fmt, hipStream_t st) {
  if (c->d_pyr_oct) CVR_HIP_CHECK(hipFree(c->d_pyr_oct));
  c->d_pyr_oct = nullptr;
  const size_t rec = fmt == 0 ? 32 : 16;
  if (hipMalloc(&c->d_pyr_oct, total * rec) != hipSuccess) {
    c->d_pyr_oct = nullptr;
    set_error("pyramid: out of device memory (corner blocks)");
    return CVR_ERR_NOMEM;
  }
  const PyrGeom& pg = c->pyr;
  for (int l = 0; l < pg.levels; ++l) {
    const int nx = pg.dims[l][0], ny = pg.dims[l][1], nz = pg.dims[l][2];
    if (nz > 65535) { set_error("corner blocks: depth above 65535"); return CVR_ERR_INVALID; }
    const dim3 grid(unsigned((nx + 63) / 64), unsigned((ny + 3) / 4), unsigned(nz));
    void* d = static_cast<char*>(c->d_pyr_oct) + pg.boff[l] * rec;
    const float* s = c->d_pyr + pg.off[l];
    if (fmt == 0) hipLaunchKernelGGL(k_make_oct<0>, grid, dim3(256), 0, st, s, d, nx, ny, nz);
    else if (fmt == 1) hipLaunchKernelGGL(k_make_oct<1>, grid, dim3(256), 0, st, s, d, nx, ny, nz);
    else hipLaunchKernelGGL(k_make_oct<2>, grid, dim3(256), 0, st, s, d, nx, ny, nz);
    CVR_HIP_CHECK(hipGetLastError());
  }
  CVR_HIP_CHECK(hipStreamSynchronize(st));
  return CVR_OK;
}

}  // namespace cvr
