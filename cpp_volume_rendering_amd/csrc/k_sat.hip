// 3D summed-area table of extinction on the GPU, gfx950.
//
// Replaces the CPU build of the reference (ebsrenderer.cpp:624-723 +
// vis_utils/summedareatable.h:218-278): a (nx+2)(ny+2)(nz+2) grid, zero border,
// interior GetExtN(normalised voxel), inclusive prefix sums along x, y and z.
// The prefix sums accumulate in fp32 as two-float (hi, lo) pairs (Knuth TwoSum),
// so the table keeps ~2^-44 relative accuracy while every add is an fp32 add; the
// cell is stored as float2 (hi, lo).
//
// Passes 1 and 2 (x and y) run as one kernel (k_sat_xy) when a row has at most 1024
// cells: one workgroup per z-slice, the voxels read twice and the table written once.
// Otherwise:
// Pass 1 (x): one wavefront per row, a 64-lane shuffle scan per 64-cell chunk
//   with the chunk carry broadcast from lane 63; the voxels (and their extinction
//   through the per-value LUT) of up to kXChunks chunks are loaded before the chunks
//   are scanned in turn; writes float2.
// Pass 2 (y) / pass 3 (z): segmented column scans.  A workgroup takes 64 columns,
//   lanes along x (coalesced); each of its wavefronts holds one segment of every
//   column in registers (all of its loads in flight at once), scans it, and adds the
//   totals of the segments before it, exchanged through LDS: every cell is read and
//   written once.  Columns longer than 16 segments of kSegCells fall back to a lane
//   walking its whole column.
#include "cvr_ctx.hpp"

namespace cvr {

struct df {
  float hi, lo;
};
__device__ __forceinline__ df two_sum(float a, float b) {
  float s = a + b;
  float bb = s - a;
  float e = (a - (s - bb)) + (b - bb);
  return df{s, e};
}
__device__ __forceinline__ df df_add(df a, df b) {
  df s = two_sum(a.hi, b.hi);
  float e = s.lo + (a.lo + b.lo);
  float hi = s.hi + e;
  float lo = e - (hi - s.hi);
  return df{hi, lo};
}

constexpr int kXChunks = 8;    // 64-cell chunks of a row whose loads are issued together
constexpr int kSegCells = 48;  // cells of a column segment held in registers

template <typename T>
__global__ __launch_bounds__(256) void k_sat_x(const T* __restrict__ vox, const float* __restrict__ lut, int nx,
                                                int ny, int nz, float2* __restrict__ sat) {
  const int w = nx + 2, h = ny + 2, d = nz + 2;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= h * d) return;
  const int y = row % h, z = row / h;
  const bool border_row = (y == 0 || z == 0 || y == h - 1 || z == d - 1);
  const T* src = border_row ? vox : vox + (size_t(z - 1) * ny + (y - 1)) * nx;
  float2* dstrow = sat + size_t(row) * w;
  df carry{0.f, 0.f};
  for (int base0 = 0; base0 < w; base0 += 64 * kXChunks) {
    float vv[kXChunks];
#pragma unroll
    for (int c = 0; c < kXChunks; ++c) {
      const int x = base0 + 64 * c + lane;
      vv[c] = (!border_row && x >= 1 && x <= nx) ? lut[src[x - 1]] : 0.f;
    }
#pragma unroll
    for (int c = 0; c < kXChunks; ++c) {
      const int base = base0 + 64 * c;
      if (base >= w) break;
      const int x = base + lane;
      df acc{vv[c], 0.f};
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        float oh = __shfl_up(acc.hi, off, 64);
        float ol = __shfl_up(acc.lo, off, 64);
        if (lane >= off) acc = df_add(acc, df{oh, ol});
      }
      acc = df_add(acc, carry);
      if (x < w) dstrow[x] = make_float2(acc.hi, acc.lo);
      carry.hi = __shfl(acc.hi, 63, 64);
      carry.lo = __shfl(acc.lo, 63, 64);
    }
  }
}

// Passes x and y fused, for rows of up to 64 E cells: one workgroup per z-slice of
// xy_waves<E>() wavefronts, wavefront s taking rows [s R, (s + 1) R), lane l the cells
// l + 64 k (k < E) of each, so every load and store is coalesced.
//  - Phase 1: every wavefront sums its rows' extinction per column (its segment's
//    column totals, exchanged through LDS).
//  - Phase 2: a lane's running column sums start at the totals of the segments above;
//    each row adds its extinction to them, and the row's prefix along x is a DPP scan of
//    each 64-cell chunk plus the total of the chunks before it (lane 63's value).
// The voxels are read twice, the second time from L2 (gathering the LUT again measured
// faster than parking the extinction in the table between the phases); the table is
// written once.  The sums run in fp64: one add per cell and direction, where a two-float
// add takes about ten fp32 operations.  A cell stores hi = fp32(v), lo = fp32(v - hi).
// Workgroups 0 .. nz - 1 take the interior slices z = 1 .. nz and the last two the zero
// border slices (stores only), so a grid of nz + 2 <= CUs + 2 workgroups leaves no
// interior slice for a second round.
template <int E>
constexpr int xy_waves() {
  return E <= 6 ? 16 : 8;  // 16 wavefronts hold at most 128 VGPRs each
}

// A lane's 64-bit value moved by a DPP control (kCtrl) within the rows kRowMask
// enables; +0 in lanes with no source and in the rows left out.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ double dpp0(double v) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  const unsigned lo = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(b)), kCtrl, kRowMask, 0xf, true));
  const unsigned hi = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(b >> 32)), kCtrl, kRowMask, 0xf, true));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
// Inclusive prefix sum across the wavefront: shifts by 1, 2, 4, 8 within each 16-lane
// row, then lane 15 of rows 0 / 2 into rows 1 / 3 and lane 31 into rows 2 and 3.
__device__ __forceinline__ double wave_scan(double x) {
  x += dpp0<0x111, 0xf>(x);  // row_shr:1
  x += dpp0<0x112, 0xf>(x);  // row_shr:2
  x += dpp0<0x114, 0xf>(x);  // row_shr:4
  x += dpp0<0x118, 0xf>(x);  // row_shr:8
  x += dpp0<0x142, 0xa>(x);  // row_bcast:15
  x += dpp0<0x143, 0xc>(x);  // row_bcast:31
  return x;
}
// Lane 63's value, in every lane.
__device__ __forceinline__ double lane63(double v) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(unsigned(b)), 63));
  const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(unsigned(b >> 32)), 63));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

template <typename T, int E>
__global__ __launch_bounds__(64 * xy_waves<E>()) void k_sat_xy(const T* __restrict__ vox,
                                                                const float* __restrict__ lut, int nx, int ny, int nz,
                                                                float2* __restrict__ sat) {
  constexpr int kXyWaves = xy_waves<E>();
  constexpr int G = E <= 8 ? 4 : 2;  // rows whose loads are issued together
  __shared__ double tot[kXyWaves][64 * E];
  const int w = nx + 2, h = ny + 2;
  const int lane = threadIdx.x & 63, sg = threadIdx.x >> 6;
  const int R = (h + kXyWaves - 1) / kXyWaves;
  const int ya = min(h, sg * R), yb = min(h, ya + R);
  const int bz = int(blockIdx.x);
  const int z = bz < nz ? bz + 1 : (bz == nz ? 0 : nz + 1);
  float2* slice = sat + size_t(z) * w * h;
  if (bz >= nz) {  // a border slice: zeros
    for (int y = ya; y < yb; ++y)
#pragma unroll
      for (int k = 0; k < E; ++k)
        if (lane + 64 * k < w) slice[size_t(y) * w + lane + 64 * k] = make_float2(0.f, 0.f);
    return;
  }
  // voxel values of rows y0 .. y0 + G - 1 at this lane's cells; -1 on the border and
  // past the segment (nothing is read there)
  auto raw = [&](int y0, int (&q)[G][E]) {
#pragma unroll
    for (int r = 0; r < G; ++r) {
      const int y = y0 + r;
      const bool in = y < yb && y > 0 && y < h - 1;
      const T* src = vox + (size_t(z - 1) * ny + (in ? y - 1 : 0)) * nx;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const int x = lane + 64 * k;
        q[r][k] = (in && x >= 1 && x <= nx) ? int(src[x - 1]) : -1;
      }
    }
  };
  auto ext = [&](const int (&q)[G][E], float (&e)[G][E]) {
#pragma unroll
    for (int r = 0; r < G; ++r)
#pragma unroll
      for (int k = 0; k < E; ++k) e[r][k] = q[r][k] >= 0 ? lut[q[r][k]] : 0.f;
  };
  double c[E];
#pragma unroll
  for (int k = 0; k < E; ++k) c[k] = 0.0;
  int qn[G][E];
  raw(ya, qn);
  for (int y = ya; y < yb; y += G) {
    float e[G][E];
    ext(qn, e);
    raw(y + G, qn);
#pragma unroll
    for (int r = 0; r < G; ++r)
#pragma unroll
      for (int k = 0; k < E; ++k) c[k] += double(e[r][k]);
  }
#pragma unroll
  for (int k = 0; k < E; ++k) tot[sg][lane + 64 * k] = c[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < E; ++k) {
    c[k] = 0.0;
    for (int t = 0; t < sg; ++t) c[k] += tot[t][lane + 64 * k];
  }
  raw(ya, qn);
  for (int y = ya; y < yb; y += G) {
    float e[G][E];
    ext(qn, e);
    raw(y + G, qn);  // the next rows' voxels in flight while these rows are scanned
#pragma unroll
    for (int r = 0; r < G; ++r) {
      if (y + r >= yb) break;
      float2* row = slice + size_t(y + r) * w;
      double carry = 0.0;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        c[k] += double(e[r][k]);
        const double v = wave_scan(c[k]) + carry;
        const float hi = float(v);
        if (lane + 64 * k < w) row[lane + 64 * k] = make_float2(hi, float(v - double(hi)));
        if (k + 1 < E) carry = lane63(v);
      }
    }
  }
}

// Segmented prefix along a strided axis: lines enumerated by (i0 in [0, n0), i1 in
// [0, n1)) with base offset i0 + i1 * stride1, `n_line` cells `stride` apart.  Lane l of
// every wavefront takes line 64 blockIdx.x + l; wavefront s takes cells [s seg, (s + 1)
// seg) of it, seg = ceil(n_line / NSEG) <= kSegCells, in registers.
template <int NSEG>
__global__ __launch_bounds__(64 * NSEG) void k_sat_seg(float2* __restrict__ sat, int n0, int n1, size_t stride1,
                                                        int n_line, size_t stride) {
  __shared__ float2 tot[NSEG][64];
  const int lane = threadIdx.x & 63, sg = threadIdx.x >> 6;
  const int i0 = blockIdx.x * 64 + lane, i1 = blockIdx.y;
  const bool line_ok = i0 < n0 && i1 < n1;
  const int seg = (n_line + NSEG - 1) / NSEG;
  const int k0 = sg * seg, nk = line_ok ? min(seg, n_line - k0) : 0;  // cells of this segment (may be <= 0)
  float2* p = sat + size_t(line_ok ? i0 : 0) + size_t(line_ok ? i1 : 0) * stride1 + size_t(k0) * stride;
  df v[kSegCells];
#pragma unroll
  for (int j = 0; j < kSegCells; ++j) {
    float2 c = make_float2(0.f, 0.f);
    if (j < nk) c = p[size_t(j) * stride];
    v[j] = df{c.x, c.y};
  }
#pragma unroll
  for (int j = 1; j < kSegCells; ++j)
    if (j < nk) v[j] = df_add(v[j - 1], v[j]);
  // the segment's total (its last cell), then the carry of the segments before it
  df last{0.f, 0.f};
#pragma unroll
  for (int j = 0; j < kSegCells; ++j)
    if (j == nk - 1) last = v[j];
  tot[sg][lane] = make_float2(last.hi, last.lo);
  __syncthreads();
  df carry{0.f, 0.f};
  for (int t = 0; t < sg; ++t) carry = df_add(carry, df{tot[t][lane].x, tot[t][lane].y});
#pragma unroll
  for (int j = 0; j < kSegCells; ++j) {
    if (j < nk) {
      const df r = df_add(carry, v[j]);
      p[size_t(j) * stride] = make_float2(r.hi, r.lo);
    }
  }
}

// Prefix along a strided axis: `n_line` cells per line, `stride` between them,
// lines enumerated by (i0 in [0, n0), i1 in [0, n1)) with base offset i0 + i1*stride1.
__global__ __launch_bounds__(256) void k_sat_strided(float2* __restrict__ sat, int n0, int n1, size_t stride1,
                                                      int n_line, size_t stride) {
  const int i0 = blockIdx.x * 256 + threadIdx.x;
  const int i1 = blockIdx.y;
  if (i0 >= n0 || i1 >= n1) return;
  float2* p = sat + size_t(i0) + size_t(i1) * stride1;
  df acc{0.f, 0.f};
  int k = 0;
  for (; k + 8 <= n_line; k += 8) {
    float2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[size_t(k + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      acc = df_add(acc, df{v[u].x, v[u].y});
      p[size_t(k + u) * stride] = make_float2(acc.hi, acc.lo);
    }
  }
  for (; k < n_line; ++k) {
    float2 v = p[size_t(k) * stride];
    acc = df_add(acc, df{v.x, v.y});
    p[size_t(k) * stride] = make_float2(acc.hi, acc.lo);
  }
}

// Per raw value v: GetExtN(v / vmax) exactly as TransferFunction1D computes it on the
// CPU (transferfunction1d.cpp:132-157, 189-197; transferfunction.h:79-82): lerp of the
// FP64 table at v/vmax * max_density, rounded to float, then log(1/(1-alpha)) in
// double, rounded to float.
static void build_ext_lut(const cvr_ctx* c, std::vector<float>& lut) {
  const int nvals = c->dtype == CVR_U8 ? 256 : 65536;
  const double vmax = c->dtype == CVR_U8 ? 255.0 : 65535.0;
  lut.resize(nvals);
  const int md = c->tf_max_density;
  for (int v = 0; v < nvals; ++v) {
    double value = (double(v) / vmax) * (double(md) / 1.0);
    float a;
    if (value < 0.0f || value > float(md)) {
      a = 0.f;
    } else if (std::fabs(value - float(md)) < 0.000001) {
      a = float(c->tf_tab[size_t(md) * 4 + 3]);
    } else {
      int i = int(value);
      double t = value - i;
      a = float((1.0 - t) * c->tf_tab[size_t(i) * 4 + 3] + t * c->tf_tab[size_t(i + 1) * 4 + 3]);
    }
    lut[v] = c->tf_ext_input ? a : float(std::log(1.0 / (1.0 - double(a))));
  }
}

// Corner blocks of the SAT: record (x, y, z) holds the 8 cells of (x..x+1, y..y+1,
// z..z+1), the +1 neighbours clamped to the grid, as 4 float4 (two (hi, lo) cells each):
// planes z and z+1, each (x, y), (x+1, y) then (x, y+1), (x+1, y+1).  64 bytes, one
// cache line: a GL_LINEAR tap of the SAT is one record instead of 8 scattered reads.
// Four lanes per record, lane p writing float4 p (bit 0: y + 1, bit 1: z + 1), so a
// wavefront's store covers 16 whole records (1 KiB contiguous) instead of one quarter of
// each of 64 records; the stores are non-temporal, as the copy (64 bytes per cell) is far
// larger than the caches (256^3: 0.41 ms one lane per record, 0.33 ms whole records,
// 0.25 ms non-temporal).
__global__ __launch_bounds__(256) void k_sat_oct(const float2* __restrict__ sat, float4* __restrict__ oct, int w, int h,
                                                 int d) {
  const int x = blockIdx.x * 64 + (threadIdx.x >> 2), part = threadIdx.x & 3;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= w || y >= h || z >= d) return;
  const int x1 = min(x + 1, w - 1);
  const int yy = (part & 1) ? min(y + 1, h - 1) : y;
  const int zz = (part & 2) ? min(z + 1, d - 1) : z;
  const size_t r = (size_t(zz) * h + yy) * w;
  const float2 a = sat[r + x], b = sat[r + x1];
  typedef float nf4 __attribute__((ext_vector_type(4)));
  __builtin_nontemporal_store(nf4{a.x, a.y, b.x, b.y},
                              reinterpret_cast<nf4*>(oct) + 4 * ((size_t(z) * h + y) * w + x) + part);
}

// Corner blocks of a built SAT (64 bytes per cell), for the marches that read them.
static int build_sat_oct(cvr_ctx* c, hipStream_t st) {
  const int w = c->nx + 2, h = c->ny + 2, d = c->nz + 2;
  const size_t cells = size_t(w) * h * d;
  if (!c->d_sat_oct) {
    if (hipMalloc(&c->d_sat_oct, cells * 4 * sizeof(float4)) != hipSuccess) {
      c->d_sat_oct = nullptr;
      set_error("cvr_prepare: out of device memory for the summed-area table's corner blocks");
      return -5;
    }
  }
  if (d > 65535 || h > 65535) { set_error("summed-area table: height or depth above 65535"); return -4; }
  hipLaunchKernelGGL(k_sat_oct, dim3(unsigned((w + 63) / 64), unsigned(h), unsigned(d)), dim3(256), 0, st,
                     c->d_sat, static_cast<float4*>(c->d_sat_oct), w, h, d);
  CVR_HIP_CHECK(hipGetLastError());
  c->sat_oct_built_for = c->sat_built_for;
  return 0;
}

// A prefix pass along a strided axis: segmented (the fewest power-of-two segments of at
// most kSegCells cells) up to 16 segments, else one lane per line.
static int scan_lines(float2* sat, int n0, int n1, size_t stride1, int n_line, size_t stride, hipStream_t st) {
  const dim3 g(unsigned((n0 + 63) / 64), unsigned(n1));
  int nseg = 1;
  while (nseg * kSegCells < n_line && nseg < 16) nseg *= 2;
  if (nseg * kSegCells < n_line) {
    hipLaunchKernelGGL(k_sat_strided, dim3((n0 + 255) / 256, n1), dim3(256), 0, st, sat, n0, n1, stride1, n_line,
                       stride);
  } else {
    auto go = [&](auto ns) {
      constexpr int NS = decltype(ns)::value;
      hipLaunchKernelGGL(k_sat_seg<NS>, g, dim3(64 * NS), 0, st, sat, n0, n1, stride1, n_line, stride);
    };
    switch (nseg) {
      case 1: go(std::integral_constant<int, 1>{}); break;
      case 2: go(std::integral_constant<int, 2>{}); break;
      case 4: go(std::integral_constant<int, 4>{}); break;
      case 8: go(std::integral_constant<int, 8>{}); break;
      default: go(std::integral_constant<int, 16>{}); break;
    }
  }
  CVR_HIP_CHECK(hipGetLastError());
  return 0;
}

// Passes x and y in one launch (k_sat_xy) when a row fits 16 cells per lane; false
// otherwise (the caller runs them as two passes).
static bool launch_sat_xy(cvr_ctx* c, hipStream_t st) {
  const int w = c->nx + 2, d = c->nz + 2;
  const int need = (w + 63) / 64;
  bool launched = false;
  auto go = [&](auto ecst) {
    constexpr int E = decltype(ecst)::value;
    if (launched || need > E) return;
    launched = true;
    if (c->dtype == CVR_U8)
      hipLaunchKernelGGL((k_sat_xy<uint8_t, E>), dim3(unsigned(d)), dim3(64 * xy_waves<E>()), 0, st,
                         static_cast<const uint8_t*>(c->d_vol), c->d_ext_lut, c->nx, c->ny, c->nz, c->d_sat);
    else
      hipLaunchKernelGGL((k_sat_xy<uint16_t, E>), dim3(unsigned(d)), dim3(64 * xy_waves<E>()), 0, st,
                         static_cast<const uint16_t*>(c->d_vol), c->d_ext_lut, c->nx, c->ny, c->nz, c->d_sat);
  };
  go(std::integral_constant<int, 2>{});
  go(std::integral_constant<int, 3>{});
  go(std::integral_constant<int, 4>{});
  go(std::integral_constant<int, 5>{});
  go(std::integral_constant<int, 6>{});
  go(std::integral_constant<int, 8>{});
  go(std::integral_constant<int, 12>{});
  go(std::integral_constant<int, 16>{});
  return launched;
}

int build_sat(cvr_ctx* c, hipStream_t st, bool with_blocks) {
  const uint64_t key = (c->vol_version << 32) ^ c->tf_version;
  if (c->d_sat && c->sat_built_for == key) {
    if (with_blocks && !(c->d_sat_oct && c->sat_oct_built_for == key)) {
      if (const int rc = build_sat_oct(c, st)) return rc;
      CVR_HIP_CHECK(hipStreamSynchronize(st));
    }
    return 0;
  }
  const int w = c->nx + 2, h = c->ny + 2, d = c->nz + 2;
  const size_t cells = size_t(w) * h * d;
  if (!c->d_sat) {
    if (hipMalloc(&c->d_sat, cells * sizeof(float2)) != hipSuccess) {
      c->d_sat = nullptr;
      set_error("cvr_prepare: out of device memory for the summed-area table");
      return -5;
    }
  }
  std::vector<float> lut;
  build_ext_lut(c, lut);
  if (c->d_ext_lut) CVR_HIP_CHECK(hipFree(c->d_ext_lut));
  c->d_ext_lut = nullptr;
  CVR_HIP_CHECK(hipMalloc(&c->d_ext_lut, lut.size() * sizeof(float)));
  CVR_HIP_CHECK(hipMemcpyAsync(c->d_ext_lut, lut.data(), lut.size() * sizeof(float), hipMemcpyHostToDevice, st));
  const bool timed = c->collect_stats;
  if (timed) {
    for (hipEvent_t* e : {&c->ev_sat[0], &c->ev_sat[1], &c->ev_sat[2]})
      if (!*e) CVR_HIP_CHECK(hipEventCreate(e));
    CVR_HIP_CHECK(hipEventRecord(c->ev_sat[0], st));
  }
  if (launch_sat_xy(c, st)) {
    CVR_HIP_CHECK(hipGetLastError());
  } else {
    const int rows = h * d;
    if (c->dtype == CVR_U8)
      hipLaunchKernelGGL(k_sat_x<uint8_t>, dim3((rows + 3) / 4), dim3(256), 0, st,
                         static_cast<const uint8_t*>(c->d_vol), c->d_ext_lut, c->nx, c->ny, c->nz, c->d_sat);
    else
      hipLaunchKernelGGL(k_sat_x<uint16_t>, dim3((rows + 3) / 4), dim3(256), 0, st,
                         static_cast<const uint16_t*>(c->d_vol), c->d_ext_lut, c->nx, c->ny, c->nz, c->d_sat);
    CVR_HIP_CHECK(hipGetLastError());
    // y: lines over (x, z), stride w
    if (const int rc = scan_lines(c->d_sat, w, d, size_t(w) * h, h, size_t(w), st)) return rc;
  }
  // z: lines over (x, y), stride w*h
  if (const int rc = scan_lines(c->d_sat, w, h, size_t(w), d, size_t(w) * h, st)) return rc;
  if (timed) CVR_HIP_CHECK(hipEventRecord(c->ev_sat[1], st));
  c->sat_built_for = key;
  c->sat_oct_built_for = ~0ull;
  // corner blocks for the marches that read them
  if (with_blocks)
    if (const int rc = build_sat_oct(c, st)) return rc;
  if (timed) CVR_HIP_CHECK(hipEventRecord(c->ev_sat[2], st));
  // the LUT is read by pass 1 only; keep it until the stream has consumed it
  CVR_HIP_CHECK(hipStreamSynchronize(st));
  if (timed) {
    float a = 0.f, b = 0.f;
    CVR_HIP_CHECK(hipEventElapsedTime(&a, c->ev_sat[0], c->ev_sat[1]));
    CVR_HIP_CHECK(hipEventElapsedTime(&b, c->ev_sat[1], c->ev_sat[2]));
    c->stats.sat_scan_ms = a;
    c->stats.sat_blocks_ms = with_blocks ? b : 0.0;
  }
  return 0;
}

}  // namespace cvr
