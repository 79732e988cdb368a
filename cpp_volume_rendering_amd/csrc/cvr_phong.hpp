// Gradient-based Blinn-Phong shading shared by the plain, EBS, DOS and light-cache
// marchers: the gradient grid lookup and the two dot products of ShadeBlinnPhong
// (rc1pass/ray_marching_1p.comp:48-81; the Phong branches of
// rc1pextbsd/ebs_ray_bbox_marching.comp:523-546, rc1pdosct/ray_bbox_marching.comp:627-650,
// _common_shaders/obj_ray_marching.comp:186-211).
#pragma once
#include "cvr_common.hpp"

namespace cvr {

struct PhongK {
  int on;            // shade with the gradient (the gradient grid is bound)
  float ks, shininess;
  float spec[3];     // light source specular colour
  float lpos[3];     // WorldLightingPos / LightSourcePosition
};

// GL_LINEAR / CLAMP_TO_EDGE lookup of the gradient grid (float4 per voxel, brick order
// like the neighbour blocks) at texture-space position p.
__device__ __forceinline__ f3 sample_gradient(const float4* __restrict__ gr, const VolGeom& g, f3 p) {
  int x0, x1, y0, y1, z0, z1;
  float fx, fy, fz;
  lin_axis(p.x * g.tex_mul[0] - 0.5f, g.nx, x0, x1, fx);
  lin_axis(p.y * g.tex_mul[1] - 0.5f, g.ny, y0, y1, fy);
  lin_axis(p.z * g.tex_mul[2] - 0.5f, g.nz, z0, z1, fz);
  const float4 c000 = gr[brick_index(x0, y0, z0, g.nbx, g.nby)], c100 = gr[brick_index(x1, y0, z0, g.nbx, g.nby)];
  const float4 c010 = gr[brick_index(x0, y1, z0, g.nbx, g.nby)], c110 = gr[brick_index(x1, y1, z0, g.nbx, g.nby)];
  const float4 c001 = gr[brick_index(x0, y0, z1, g.nbx, g.nby)], c101 = gr[brick_index(x1, y0, z1, g.nbx, g.nby)];
  const float4 c011 = gr[brick_index(x0, y1, z1, g.nbx, g.nby)], c111 = gr[brick_index(x1, y1, z1, g.nbx, g.nby)];
  auto lerp3 = [](float4 a, float4 b, float t) {
    return mk3(a.x + t * (b.x - a.x), a.y + t * (b.y - a.y), a.z + t * (b.z - a.z));
  };
  auto mix3 = [](f3 a, f3 b, float t) { return mk3(a.x + t * (b.x - a.x), a.y + t * (b.y - a.y), a.z + t * (b.z - a.z)); };
  const f3 c0 = mix3(lerp3(c000, c100, fx), lerp3(c010, c110, fx), fy);
  const f3 c1 = mix3(lerp3(c001, c101, fx), lerp3(c011, c111, fx), fy);
  return mix3(c0, c1, fz);
}

// max(0, n.l) and max(0, h.n)^shininess at texture-space p (world position p - G/2);
// false where the interpolated gradient is the zero vector (the sample keeps its colour).
__device__ __forceinline__ bool phong_terms(const float4* __restrict__ gr, const VolGeom& g, const PhongK& P, f3 p,
                                            f3 eye, float& dd, float& sp) {
  const f3 n0 = sample_gradient(gr, g, p);
  if (n0.x == 0.f && n0.y == 0.f && n0.z == 0.f) return false;
  const f3 n = norm3(n0);
  const f3 wp = p - mk3(0.5f * g.G[0], 0.5f * g.G[1], 0.5f * g.G[2]);
  const f3 L = norm3(ld3(P.lpos) - wp);
  const f3 E = norm3(eye - wp);
  const f3 Hv = norm3(E + L);
  dd = fmaxf(0.f, dot3(n, L));
  sp = __powf(fmaxf(0.f, dot3(Hv, n)), P.shininess);
  return true;
}

}  // namespace cvr
