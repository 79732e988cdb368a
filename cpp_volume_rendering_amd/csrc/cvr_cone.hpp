// Device helpers of the cone-traced lighting: pyramid lookups, cone transmittance,
// directional occlusion / shadow, and the surrogate's input features.
#pragma once
#include "cvr_ctx.hpp"
#include <hip/hip_fp16.h>
#include <type_traits>

namespace cvr {

// One step of a cone march: midpoint distance, segment length, pyramid level.
struct ConeStepF {
  float tm, dt, lod, pad;
};
// One stencil point of the surrogate's features: distance along the axis, segment
// length, pyramid level.
struct StencilF {
  float dist, seg, lod, pad;
};

// Surrogate inputs: kNOcc occlusion segments read twice (on the axial cone, on the
// ring), kNSdw shadow segments, the local opacity.
constexpr int kNFeat = 16, kNOcc = 4, kNSdw = 7, kNStencil = 2 * kNOcc + kNSdw, kMaxSteps = 256;

// One surrogate input's trilinear read with its pyramid level resolved on the host:
// the stencil point (distance along its axis; segment length x log2(e), 1 x log2(e) for
// the local opacity) and the geometry of the level it reads (texel coordinate per world
// unit, last texel index per axis as a float, bricks, byte offset of the level's corner
// blocks, border falloff log2(e) / (2 (2^l)^2)), with log2(e) folded in so that both
// exponentials of a read are single exp2 instructions.  64 bytes; the table is read
// through scalar loads.
struct FeatTap {
  float dist, seg;
  float ts[3];
  float fall_k;
  float dmax[3];
  int nb[2];
  unsigned int boff_lo, boff_hi;
  int level, pad[2];
};
static_assert(sizeof(FeatTap) == 64, "FeatTap is 16 dwords");

// Host dispatch over the compile-time variants of the pyramid-reading kernels: calls
// fn(Tag<voxel type>{}, integral_constant<pyramid storage>{}).
template <class T> struct Tag { typedef T type; };
template <class Fn> void with_fmt(int fmt, Fn&& fn) {
  if (fmt == 1) fn(std::integral_constant<int, 1>{});
  else if (fmt == 2) fn(std::integral_constant<int, 2>{});
  else fn(std::integral_constant<int, 0>{});
}
template <class Fn> void with_types(int dtype, int fmt, Fn&& fn) {
  with_fmt(fmt, [&](auto f) {
    if (dtype == CVR_U8) fn(Tag<uint8_t>{}, f);
    else fn(Tag<uint16_t>{}, f);
  });
}

// Per-frame cone-tracer constants.
struct ConeK {
  int n_steps, n_occ, apply_occ, apply_sdw;
  float ring_cos, ring_sin, ka, kd;
  float L[3];               // shadow cone axis (towards the light)
  const ConeStepF* so;      // occlusion schedule [n_steps]
  const ConeStepF* ss;      // shadow schedule [n_steps]
  const FeatTap* ft;        // surrogate inputs: axial/ring pairs [2*kNOcc], shadow [kNSdw], local [1]
};

// One corner block of the pyramid in its storage format: the values at (x..x+1, y..y+1,
// z..z+1) of a level, the +1 neighbours clamped to the edge; fp32 as two float4 (planes
// z and z+1, each (x, y), (x+1, y), (x, y+1), (x+1, y+1)), fp16 / bf16 as eight 16-bit
// values in 16 bytes.  A trilinear read is one record: one cache line, one address.
template <int F> struct PyrOct { typedef uint4 type; };
template <> struct PyrOct<0> { struct type { float4 a, b; }; };
struct Oct8 { float4 a, b; };

template <int F>
__device__ __forceinline__ Oct8 pyr_decode(typename PyrOct<F>::type r) {
  if constexpr (F == 0) {
    return Oct8{r.a, r.b};
  } else if constexpr (F == 1) {
    auto h = [](unsigned int u, bool hi) { return __half2float(__ushort_as_half(ushort(hi ? u >> 16 : u & 0xFFFFu))); };
    return Oct8{make_float4(h(r.x, false), h(r.x, true), h(r.y, false), h(r.y, true)),
                make_float4(h(r.z, false), h(r.z, true), h(r.w, false), h(r.w, true))};
  } else {
    auto b = [](unsigned int u, bool hi) { return __uint_as_float(hi ? u & 0xFFFF0000u : u << 16); };
    return Oct8{make_float4(b(r.x, false), b(r.x, true), b(r.y, false), b(r.y, true)),
                make_float4(b(r.z, false), b(r.z, true), b(r.w, false), b(r.w, true))};
  }
}

// Trilinear value of one corner block at weights (fx, fy, fz): quad_lerp of the decoded
// record.  For fp16 records the four x lerps take the two halves of each dword (the x
// pair at one y and z) straight from the record with mixed-precision FMAs
// (v_fma_mix_f32: f16 operands, fp32 arithmetic): d = hi * 1 - lo is the fp32
// difference rounded once, and lo + fx * d the fused lerp, the same values and roundings
// as converting the eight texels first, in 8 instructions instead of 16.  `one` is 1.0f
// from a kernel argument (PyrGeom::one), opaque to the compiler, so that hi * 1 - lo
// stays one fused instruction.
typedef _Float16 half2v_t __attribute__((ext_vector_type(2)));
template <int F>
__device__ __forceinline__ float pyr_lerp(typename PyrOct<F>::type r, float fx, float fy, float fz, float one) {
  if constexpr (F == 1) {
    auto xl = [&](unsigned int w) {
      const half2v_t h = __builtin_bit_cast(half2v_t, w);
      const float lo = float(h[0]), hi = float(h[1]);
      return __builtin_fmaf(fx, __builtin_fmaf(hi, one, -lo), lo);
    };
    const float a0 = xl(r.x), a1 = xl(r.y), b0 = xl(r.z), b1 = xl(r.w);
    const float c0 = a0 + fy * (a1 - a0), c1 = b0 + fy * (b1 - b0);
    return c0 + fz * (c1 - c0);
  } else {
    const Oct8 o = pyr_decode<F>(r);
    return quad_lerp(o.a, o.b, fx, fy, fz);
  }
}

// GL_LINEAR lookup of pyramid level `l` at texture-space point tp: the texel coordinate
// tp * dims / G - 1/2 is one fma per axis, the read one corner block.
template <int F>
__device__ __forceinline__ float pyr_tri(const void* __restrict__ pyr, const PyrGeom& pg, int l, f3 tp) {
  int x0, y0, z0;
  float fx, fy, fz;
  quad_axis_f(fmaf(tp.x, pg.ts[l][0], -0.5f), pg.dmax[l][0], x0, fx);
  quad_axis_f(fmaf(tp.y, pg.ts[l][1], -0.5f), pg.dmax[l][1], y0, fy);
  quad_axis_f(fmaf(tp.z, pg.ts[l][2], -0.5f), pg.dmax[l][2], z0, fz);
  typedef typename PyrOct<F>::type B;
  return pyr_lerp<F>(static_cast<const B*>(pyr)[pg.boff[l] + brick_index(x0, y0, z0, pg.nb[l][0], pg.nb[l][1])], fx, fy,
                     fz, pg.one);
}

// Border falloff of GetGaussianExtinction (rc1pdosct/ray_bbox_marching.comp:96-109):
// exp(-|clamp(p, 0, G) - p|^2 / (2 (2^lod)^2)); the distance is 0 inside the box, so the
// factor is exactly 1 there and no branch is needed (the loads of many lookups can then
// be issued back to back).  inv2sg2 = 1 / (2 (2^lod)^2).
__device__ __forceinline__ float border_falloff(const PyrGeom& pg, f3 tp, float inv2sg2) {
  // clamp(p, 0, G) as one median-of-three (p is finite, G > 0)
  const float cx = __builtin_amdgcn_fmed3f(tp.x, 0.f, pg.G[0]) - tp.x;
  const float cy = __builtin_amdgcn_fmed3f(tp.y, 0.f, pg.G[1]) - tp.y;
  const float cz = __builtin_amdgcn_fmed3f(tp.z, 0.f, pg.G[2]) - tp.z;
  return __expf(-(cx * cx + cy * cy + cz * cz) * inv2sg2);
}

// textureLod with LINEAR_MIPMAP_LINEAR (lerp of levels floor(lod) and floor(lod)+1,
// the second clamped to the chain), then the border falloff.
template <int F>
__device__ __forceinline__ float gauss_ext(const void* __restrict__ pyr, const PyrGeom& pg, f3 tp, float lod) {
  const float q = float(pg.levels - 1);
  lod = fminf(fmaxf(lod, 0.f), q);
  const int d1 = int(floorf(lod));
  const int d2 = min(d1 + 1, pg.levels - 1);
  const float fr = lod - float(d1);
  const float a = pyr_tri<F>(pyr, pg, d1, tp);
  const float b = pyr_tri<F>(pyr, pg, d2, tp);
  return (a * (1.f - fr) + b * fr) * border_falloff(pg, tp, 0.5f * exp2f(-2.f * lod));
}

// The same lookup at a whole level `l` (one trilinear read).
template <int F>
__device__ __forceinline__ float gauss_ext_level(const void* __restrict__ pyr, const PyrGeom& pg, f3 tp, int l) {
  return pyr_tri<F>(pyr, pg, l, tp) * border_falloff(pg, tp, __uint_as_float((126u - 2u * unsigned(l)) << 23));
}

template <int F>
__device__ __forceinline__ float cone_T(const void* __restrict__ pyr, const PyrGeom& pg, const ConeStepF* __restrict__ s,
                                        int n, f3 p, f3 axis) {
  float od = 0.f;
#pragma unroll 4
  for (int j = 0; j < n; ++j) {
    const ConeStepF c = s[j];
    od += gauss_ext<F>(pyr, pg, p + axis * c.tm, c.lod) * c.dt;
  }
  return __expf(-od);
}

// Occlusion frame of a camera ray (gt_ray_marching.comp:396-399).
__device__ __forceinline__ void occ_frame(f3 dir, f3& k, f3& up, f3& right) {
  right = norm3(cross3(dir, mk3(0.f, 1.f, 0.f)));
  up = norm3(cross3(mk3(-dir.x, -dir.y, -dir.z), right));
  k = mk3(-dir.x, -dir.y, -dir.z);
}

// 1 axial cone + (n_occ-1) cones whose axes sit on a ring at ring angle from the axis,
// cosine-weighted like gt_ray_marching.comp:161-168.
template <int F>
__device__ __forceinline__ float ct_occlusion(const void* __restrict__ pyr, const PyrGeom& pg, const ConeK& K, f3 p,
                                              f3 k, f3 up, f3 right) {
  float num = cone_T<F>(pyr, pg, K.so, K.n_steps, p, k), den = 1.f;
  const int nr = K.n_occ - 1;
  for (int j = 0; j < nr; ++j) {
    float sphi, cphi;
    __sincosf(6.283185307179586f * float(j) / float(nr), &sphi, &cphi);
    const f3 ax = k * K.ring_cos + (up * cphi + right * sphi) * K.ring_sin;
    num += K.ring_cos * cone_T<F>(pyr, pg, K.so, K.n_steps, p, ax);
    den += K.ring_cos;
  }
  return num / den;
}

// The surrogate's 16 inputs (see the CPU oracle's nn_features): for occlusion segment i
// the opacity read on the axial cone (2i) and on the ring (2i+1), 7 shadow segment
// opacities along the light (8..14), the local opacity (15).  The corner-block loads of
// a group of feat_group() inputs are issued together (texel weights and falloff x segment
// length kept in registers), so one memory round trip serves the group, and the groups
// are software-pipelined: the next group's loads are in flight while the current group
// is interpolated (512^3 surrogate frame at 1024^2, fp16: 5.80 -> 5.57 ms).  Each read is
// gauss_ext_level's, with the segment length folded into the falloff factor.  Groups of 4
// keep two groups of 32-byte fp32 records within the registers of two waves per SIMD.
// The groups only order the loads: every input is computed the same way whatever their
// size.
typedef const __attribute__((address_space(4))) FeatTap* ConstTaps;

#ifndef CVR_FEAT_GROUP
#define CVR_FEAT_GROUP 0
#endif
template <int F>
__host__ __device__ constexpr int feat_group() { return CVR_FEAT_GROUP ? CVR_FEAT_GROUP : 4; }

// Byte offset of corner block (x0, y0, z0) in a level of `nbx` x `nby` bricks, records of
// 2^kShift bytes: brick number, then the 6 bits of in-brick position (brick_index).  The
// 32-bit form needs the level's records to span less than 4 GiB (PyrGeom::wide = 0); its
// products are 24-bit (v_mad_u32_u24).
template <int kShift>
__device__ __forceinline__ uint32_t oct_offset32(int x0, int y0, int z0, int nbx, int nby) {
  const uint32_t b = __umul24(__umul24(uint32_t(z0) >> 2, uint32_t(nby)) + (uint32_t(y0) >> 2), uint32_t(nbx)) +
                     (uint32_t(x0) >> 2);
  const uint32_t local = ((uint32_t(z0) & 3u) << 4) | ((uint32_t(y0) & 3u) << 2) | (uint32_t(x0) & 3u);
  return ((b << 6) | local) << kShift;
}

template <int F, bool kWide>
__device__ __forceinline__ void nn_feats_at(const void* __restrict__ pyr, const PyrGeom& pg,
                                            const FeatTap* __restrict__ taps, f3 p, f3 k, f3 L, float* f) {
  typedef typename PyrOct<F>::type B;
  constexpr int kShift = sizeof(B) == 32 ? 5 : 4;
  static_assert(sizeof(B) == (1u << kShift), "records of 16 or 32 bytes");
  constexpr int kFeatGroup = feat_group<F>();
  static_assert(kNFeat % kFeatGroup == 0, "the load groups must tile the inputs");
  // the table is wave-uniform and read-only: constant address space -> scalar loads
  const ConstTaps ft = (ConstTaps)(uintptr_t)taps;
  // one group's records, texel weights and falloff x segment length
  struct Grp {
    B b[kFeatGroup];
    float fx[kFeatGroup], fy[kFeatGroup], fz[kFeatGroup], w[kFeatGroup];
  };
  auto issue = [&](int g0, Grp& G) {
#pragma unroll
    for (int j = 0; j < kFeatGroup; ++j) {
      const int i = g0 + j;
      const f3 a = i < 2 * kNOcc ? k : L;
      const f3 tp = i == kNFeat - 1 ? p : p + a * ft[i].dist;
      int x0, y0, z0;
      quad_axis_f(fmaf(tp.x, ft[i].ts[0], -0.5f), ft[i].dmax[0], x0, G.fx[j]);
      quad_axis_f(fmaf(tp.y, ft[i].ts[1], -0.5f), ft[i].dmax[1], y0, G.fy[j]);
      quad_axis_f(fmaf(tp.z, ft[i].ts[2], -0.5f), ft[i].dmax[2], z0, G.fz[j]);
      // the level's first byte (a wave-uniform address), then the record's byte offset
      const char* lv = static_cast<const char*>(pyr) + ((size_t(ft[i].boff_hi) << 32) | ft[i].boff_lo);
      if constexpr (kWide)
        G.b[j] = *reinterpret_cast<const B*>(lv + (brick_index(x0, y0, z0, ft[i].nb[0], ft[i].nb[1]) << kShift));
      else
        G.b[j] = *reinterpret_cast<const B*>(lv + oct_offset32<kShift>(x0, y0, z0, ft[i].nb[0], ft[i].nb[1]));
      // border falloff x segment length, both scaled by log2(e) on the host
      const float cx = __builtin_amdgcn_fmed3f(tp.x, 0.f, pg.G[0]) - tp.x;
      const float cy = __builtin_amdgcn_fmed3f(tp.y, 0.f, pg.G[1]) - tp.y;
      const float cz = __builtin_amdgcn_fmed3f(tp.z, 0.f, pg.G[2]) - tp.z;
      G.w[j] = __builtin_amdgcn_exp2f(-(cx * cx + cy * cy + cz * cz) * ft[i].fall_k) * ft[i].seg;
    }
  };
  auto finish = [&](int g0, const Grp& G) {
#pragma unroll
    for (int j = 0; j < kFeatGroup; ++j)
      f[g0 + j] = 1.f - __builtin_amdgcn_exp2f(-(pyr_lerp<F>(G.b[j], G.fx[j], G.fy[j], G.fz[j], pg.one) * G.w[j]));
  };
  // software pipeline over the groups: group g + 1's loads go out before group g is
  // computed, so one group is always in flight while the other is interpolated
  constexpr int kGroups = kNFeat / kFeatGroup;
  Grp ga, gb;
  issue(0, ga);
#pragma unroll
  for (int g = 0; g < kGroups; ++g) {
    Grp& cur = (g & 1) ? gb : ga;
    Grp& nxt = (g & 1) ? ga : gb;
    if (g + 1 < kGroups) issue((g + 1) * kFeatGroup, nxt);
    // keep the machine scheduler from pulling the lerps up between the loads
    __builtin_amdgcn_sched_barrier(0);
    finish(g * kFeatGroup, cur);
  }
}

// Dispatch on the level sizes: 32-bit record offsets unless a level spans 4 GiB or more.
template <int F>
__device__ __forceinline__ void nn_feats(const void* __restrict__ pyr, const PyrGeom& pg, const FeatTap* __restrict__ taps,
                                         f3 p, f3 k, f3 L, float* f) {
  if (pg.wide) nn_feats_at<F, true>(pyr, pg, taps, p, k, L, f);
  else nn_feats_at<F, false>(pyr, pg, taps, p, k, L, f);
}

// Network-free estimate from the same inputs (the network predicts the residual):
// transmittance products along the axial cone and the ring, weighted like the exact
// tracer's cones (ring weight w_ring = (n_occ-1) cos(ring)), and along the light.
__device__ __forceinline__ void nn_base(const float* f, float w_ring, float& occ, float& sdw) {
  float ta = 1.f, tr = 1.f, ts = 1.f;
#pragma unroll
  for (int i = 0; i < kNOcc; ++i) { ta *= 1.f - f[2 * i]; tr *= 1.f - f[2 * i + 1]; }
#pragma unroll
  for (int i = 0; i < kNSdw; ++i) ts *= 1.f - f[2 * kNOcc + i];
  occ = (ta + w_ring * tr) / (1.f + w_ring);
  sdw = ts;
}

}  // namespace cvr
