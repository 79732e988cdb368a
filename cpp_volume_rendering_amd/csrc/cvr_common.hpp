// Shared host/device definitions of the HIP ray casters (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

namespace cvr {

// Error reporting: the C ABI sets a thread-local message and returns a status.
void set_error(const std::string& msg);

#define CVR_HIP_CHECK(expr)                                                                   \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      ::cvr::set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr);    \
      return -2;                                                                              \
    }                                                                                         \
  } while (0)

// Upper bound on samples per camera ray: every march loop stops there even if a
// caller passes a vanishing step size.
constexpr int kMaxSamples = 1 << 21;

// --------------------------------------------------------------------------- camera
// Per-frame ray set-up, computed on the host in double and rounded once.
// dir(px,py) = normalize(a*s + b*u + f) is the expansion of
// `vec3(Vx*tan*aspect, Vy*tan, -1) * mat3(View)` (ray_marching_1p.comp:170) for
// View = glm::lookAt(eye, center, up) = rows (s, u, -f).
struct RayFrame {
  float eye[3];
  float s[3], u[3], f[3];
  float tanfov, aspect;
  int W, H;
};

// Volume geometry shared by all marchers.
// Brick order of the neighbour-block grids: 4 x 4 x 4 cells per brick (x fastest inside
// a brick, bricks x fastest), so that samples close in any direction share cache lines.
// (The brick number fits 32 bits for every grid the library accepts: at most 2^34 cells.)
__host__ __device__ __forceinline__ size_t brick_index(int x, int y, int z, int nbx, int nby) {
  const unsigned int b = (unsigned(z >> 2) * unsigned(nby) + unsigned(y >> 2)) * unsigned(nbx) + unsigned(x >> 2);
  return (size_t(b) << 6) | size_t(((z & 3) << 4) | ((y & 3) << 2) | (x & 3));
}
__host__ __device__ __forceinline__ int bricks(int n) { return (n + 3) >> 2; }
__host__ __device__ __forceinline__ size_t brick_cells(int nx, int ny, int nz) {
  return size_t(bricks(nx)) * size_t(bricks(ny)) * size_t(bricks(nz)) * 64u;
}

struct VolGeom {
  int nx, ny, nz, dtype;
  float G[3];        // box size, world units
  float scale[3];    // voxel size
  float tex_mul[3];  // n / G: world -> texel units
  const void* vox;   // raw voxels, x fastest
  const void* oct;   // per voxel its 2 x 2 x 2 corner block, edge-clamped (8 values), brick order
  int nbx, nby;      // bricks along x and y
  float vmax[3];     // n - 1 per axis: the last texel index as a float
};

// --------------------------------------------------------------------------- device helpers
#if defined(__HIPCC__)
struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk3(float a, float b, float c) { return f3{a, b, c}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ f3 norm3(f3 v) {
  float r = 1.0f / sqrtf(dot3(v, v));
  return v * r;
}
__device__ __forceinline__ f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }

// Camera ray of pixel (px, py), ray_marching_1p.comp:164-170.
__device__ __forceinline__ f3 camera_ray(const RayFrame& rf, int px, int py) {
  float vx = ((float(px) + 0.5f) / float(rf.W)) * 2.0f - 1.0f;
  float vy = ((float(py) + 0.5f) / float(rf.H)) * 2.0f - 1.0f;
  float a = vx * rf.tanfov * rf.aspect, b = vy * rf.tanfov;
  f3 d = ld3(rf.s) * a + ld3(rf.u) * b + ld3(rf.f);
  return norm3(d);
}

// Slab test against [-G/2, G/2], ray_bbox_intersection.comp:18-52.
__device__ __forceinline__ bool ray_box(f3 o, f3 d, const float* G, float& tnear, float& tfar) {
  float ix = 1.0f / d.x, iy = 1.0f / d.y, iz = 1.0f / d.z;
  float ax = ix * (-0.5f * G[0] - o.x), bx = ix * (0.5f * G[0] - o.x);
  float ay = iy * (-0.5f * G[1] - o.y), by = iy * (0.5f * G[1] - o.y);
  float az = iz * (-0.5f * G[2] - o.z), bz = iz * (0.5f * G[2] - o.z);
  tnear = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
  tfar = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
  bool hit = tfar > tnear;
  tnear = fmaxf(tnear, 0.0f);
  return hit;
}

// GL_LINEAR / CLAMP_TO_EDGE along one axis: texel coordinate x (already u*n - 1/2).
__device__ __forceinline__ void lin_axis(float x, int n, int& i0, int& i1, float& f) {
  float fl = floorf(x);
  f = x - fl;
  int i = int(fl);
  i0 = min(max(i, 0), n - 1);
  i1 = min(max(i + 1, 0), n - 1);
}

// GL_LINEAR along one axis for a quad layout (each cell also holds its +1
// neighbour, edge-clamped): first cell i in [0, n-1] and weight f.  Below the first
// texel centre both taps clamp to cell 0, so the weight is zeroed there.
__device__ __forceinline__ void quad_axis(float x, int n, int& i, float& f) {
  const float xc = fmaxf(x, 0.f);  // below the first texel centre: cell 0 with weight 0
  // xc >= 0: its fractional part is exact (v_fract_f32 = xc - floor(xc)) and truncation
  // is the floor
  f = __builtin_amdgcn_fractf(xc);
  i = min(int(xc), n - 1);
}

// quad_axis with the last texel index nmax = n - 1 as a float: one median-of-three clamps
// x to [0, n - 1].  Past the last texel centre the block's +1 neighbour along the axis is
// the edge texel itself, so a zero weight there gives the same value as quad_axis's
// fractional one (a + f * 0 = a), and the cell is the same.
__device__ __forceinline__ void quad_axis_f(float x, float nmax, int& i, float& f) {
  const float xc = __builtin_amdgcn_fmed3f(x, 0.f, nmax);
  f = __builtin_amdgcn_fractf(xc);
  i = int(xc);
}

// Trilinear from two quad loads (planes z0 and z1): c = (v00, v10, v01, v11) over (x, y).
__device__ __forceinline__ float quad_lerp(float4 a, float4 b, float fx, float fy, float fz) {
  const float a0 = a.x + fx * (a.y - a.x), a1 = a.z + fx * (a.w - a.z);
  const float b0 = b.x + fx * (b.y - b.x), b1 = b.z + fx * (b.w - b.z);
  const float c0 = a0 + fy * (a1 - a0), c1 = b0 + fy * (b1 - b0);
  return c0 + fz * (c1 - c0);
}

// Corner blocks of the raw volume: cell (x, y, z) holds the 8 voxels of (x..x+1,
// y..y+1, z..z+1), the +1 neighbours clamped to the edge, plane z first, each plane
// (x, y), (x+1, y), (x, y+1), (x+1, y+1): 8 bytes for u8 (uint2), 16 for u16 (uint4).
template <typename T> struct VolOct;
template <> struct VolOct<uint8_t> {
  typedef uint2 type;
  __device__ static void f(uint2 r, float4& a, float4& b) {
    a = make_float4(float(r.x & 0xFFu), float((r.x >> 8) & 0xFFu), float((r.x >> 16) & 0xFFu), float(r.x >> 24));
    b = make_float4(float(r.y & 0xFFu), float((r.y >> 8) & 0xFFu), float((r.y >> 16) & 0xFFu), float(r.y >> 24));
  }
};
template <> struct VolOct<uint16_t> {
  typedef uint4 type;
  __device__ static void f(uint4 r, float4& a, float4& b) {
    a = make_float4(float(r.x & 0xFFFFu), float(r.x >> 16), float(r.y & 0xFFFFu), float(r.y >> 16));
    b = make_float4(float(r.z & 0xFFFFu), float(r.z >> 16), float(r.w & 0xFFFFu), float(r.w >> 16));
  }
};

// Trilinear sample through the corner blocks: 1 load instead of 8, same result as
// sample_volume (same lerp order, same clamped taps).
template <typename T>
__device__ __forceinline__ float sample_volume_oct(const VolGeom& g, f3 p) {
  typedef typename VolOct<T>::type RT;
  const RT* __restrict__ q = static_cast<const RT*>(g.oct);
  int x0, y0, z0;
  float fx, fy, fz;
  quad_axis_f(p.x * g.tex_mul[0] - 0.5f, g.vmax[0], x0, fx);
  quad_axis_f(p.y * g.tex_mul[1] - 0.5f, g.vmax[1], y0, fy);
  quad_axis_f(p.z * g.tex_mul[2] - 0.5f, g.vmax[2], z0, fz);
  float4 a, b;
  VolOct<T>::f(q[brick_index(x0, y0, z0, g.nbx, g.nby)], a, b);
  return quad_lerp(a, b, fx, fy, fz) * (sizeof(T) == 1 ? (1.0f / 255.0f) : (1.0f / 65535.0f));
}

// The same read split in two, for marches that issue the loads of later samples early:
// the corner block holding the taps of p, then the lerp of that block at p.
template <typename T>
__device__ __forceinline__ typename VolOct<T>::type volume_oct_load(const VolGeom& g, f3 p) {
  typedef typename VolOct<T>::type RT;
  int x0, y0, z0;
  float fx, fy, fz;
  quad_axis_f(p.x * g.tex_mul[0] - 0.5f, g.vmax[0], x0, fx);
  quad_axis_f(p.y * g.tex_mul[1] - 0.5f, g.vmax[1], y0, fy);
  quad_axis_f(p.z * g.tex_mul[2] - 0.5f, g.vmax[2], z0, fz);
  return static_cast<const RT*>(g.oct)[brick_index(x0, y0, z0, g.nbx, g.nby)];
}
template <typename T>
__device__ __forceinline__ float volume_oct_lerp(const VolGeom& g, typename VolOct<T>::type r, f3 p) {
  int x0, y0, z0;
  float fx, fy, fz;
  quad_axis_f(p.x * g.tex_mul[0] - 0.5f, g.vmax[0], x0, fx);
  quad_axis_f(p.y * g.tex_mul[1] - 0.5f, g.vmax[1], y0, fy);
  quad_axis_f(p.z * g.tex_mul[2] - 0.5f, g.vmax[2], z0, fz);
  float4 a, b;
  VolOct<T>::f(r, a, b);
  return quad_lerp(a, b, fx, fy, fz) * (sizeof(T) == 1 ? (1.0f / 255.0f) : (1.0f / 65535.0f));
}

// Trilinear sample of the raw volume at world-in-box position p (texture space
// [0, G]); returns the normalised density v/255 or v/65535
// (structuredgridvolume.cpp:121-151 + GL_LINEAR, utils.cpp:8-9).
template <typename T>
__device__ __forceinline__ float sample_volume(const T* __restrict__ v, const VolGeom& g, f3 p) {
  int x0, x1, y0, y1, z0, z1;
  float fx, fy, fz;
  lin_axis(p.x * g.tex_mul[0] - 0.5f, g.nx, x0, x1, fx);
  lin_axis(p.y * g.tex_mul[1] - 0.5f, g.ny, y0, y1, fy);
  lin_axis(p.z * g.tex_mul[2] - 0.5f, g.nz, z0, z1, fz);
  const size_t sy = size_t(g.nx), sz = size_t(g.nx) * g.ny;
  const size_t r00 = y0 * sy + z0 * sz, r10 = y1 * sy + z0 * sz, r01 = y0 * sy + z1 * sz, r11 = y1 * sy + z1 * sz;
  float c000 = float(v[r00 + x0]), c100 = float(v[r00 + x1]);
  float c010 = float(v[r10 + x0]), c110 = float(v[r10 + x1]);
  float c001 = float(v[r01 + x0]), c101 = float(v[r01 + x1]);
  float c011 = float(v[r11 + x0]), c111 = float(v[r11 + x1]);
  float c00 = c000 + fx * (c100 - c000), c10 = c010 + fx * (c110 - c010);
  float c01 = c001 + fx * (c101 - c001), c11 = c011 + fx * (c111 - c011);
  float c0 = c00 + fy * (c10 - c00), c1 = c01 + fy * (c11 - c01);
  float c = c0 + fz * (c1 - c0);
  return c * (sizeof(T) == 1 ? (1.0f / 255.0f) : (1.0f / 65535.0f));
}

// texture(TexTransferFunc, d) over a float4 table of T entries (GL_LINEAR, CLAMP_TO_EDGE).
__device__ __forceinline__ float4 sample_tf(const float4* __restrict__ tf, int T, float d) {
  int i0, i1;
  float f;
  lin_axis(d * float(T) - 0.5f, T, i0, i1, f);
  float4 a = tf[i0], b = tf[i1];
  return make_float4(a.x + f * (b.x - a.x), a.y + f * (b.y - a.y), a.z + f * (b.z - a.z), a.w + f * (b.w - a.w));
}

// XCD-aware block order.  The hardware deals workgroups round-robin to the 8 XCDs
// (workgroup b to XCD b % 8), and each XCD has its own L2 while the MALL is shared.
// Within every run of 512 workgroups, XCD x's k-th workgroup gets index
// 512 (k / 64) + 64 x + k % 64: each XCD works through whole 64-index groups (one
// supertile of super_tile below, sized for the 64 workgroups an XCD holds at once),
// and at any moment the 8 XCDs hold 8 consecutive supertiles, i.e. one compact region
// of the image whose part of the volume the shared MALL then serves to all of them.
// (Giving each XCD a contiguous eighth of the image instead was 22 % slower on the
// 512^3 surrogate frame: 8 distant regions share nothing in the MALL.)  Workgroups past
// the last full run of 512 keep their index; a permutation for every grid size.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  if (bid >= (nblocks >> 9) << 9) return bid;
  const int x = bid & 7, k = bid >> 3;
  return ((k >> 6) << 9) | (x << 6) | (k & 63);
}

// Workgroup index -> tile of a tiles_x x tiles_y grid of tile_px-wide tiles, in
// supertiles of 64 tiles that are strip_px pixels wide: SW = strip_px / tile_px tiles
// across (at least 1) and 64 / SW down (row-major supertiles, row-major tiles inside;
// the supertiles at the right and bottom edges are narrower or shorter).  After
// xcd_remap, the 64 workgroups an XCD holds at once render one tall strip, and the 8
// XCDs 8 strips side by side.  The strip width is the caller's, from measurements on
// the 1024^2 frames: 16 pixels for the surrogate, DOS and the cone tracers (against 32:
// 5 %, 15 % and 15 % faster) and the batched surrogate (2 %), 32 for EBS and the plain
// march (against 16: 7 % and 2 % faster); 64- or 128-pixel strips were slower everywhere
// (up to 40 % for the cone tracer).  A bijection onto the grid for every size.
__host__ __device__ __forceinline__ void super_tile(int bid, int tiles_x, int tiles_y, int tile_px, int strip_px, int& tx,
                                                    int& ty) {
  const int SW = strip_px > tile_px ? strip_px / tile_px : 1, SH = 64 / SW;
  const int full_rows = tiles_y / SH, rowsz = tiles_x * SH;
  int sy, r, h;
  if (bid < full_rows * rowsz) {
    sy = bid / rowsz;
    r = bid - sy * rowsz;
    h = SH;
  } else {
    sy = full_rows;
    r = bid - full_rows * rowsz;
    h = tiles_y - full_rows * SH;
  }
  const int sx = r / (SW * h);
  const int w = sx < tiles_x / SW ? SW : tiles_x - (tiles_x / SW) * SW;
  const int q = r - sx * SW * h;
  tx = sx * SW + q % w;
  ty = sy * SH + q / w;
}
#endif

}  // namespace cvr
