// Extinction-based shading (Schlegel et al. 2011), gfx950.
//
// Semantics follow cppvolrend/structured/rc1pextbsd/ebs_ray_bbox_marching.comp:
//   ambient occlusion from nested SAT boxes (:102-146), a dominant-axis cone
//   shadow built from a chain of SAT boxes (:148-456), shading :500-551, the
//   march :553-627.
// The SAT is sampled with GL_LINEAR semantics like the reference's R32F texture,
// but every tap and every 8-tap box sum is evaluated in FP64 on the two-float
// (hi + lo) cells: the prefix values reach ~1e7 while small AO boxes hold ~10,
// and an fp32 box sum would keep no correct digit there.
// For a directional light the cone geometry is the same for every sample: each
// workgroup builds the chain of boxes once in FP64 (offsets, next-position cut
// coordinate, box volume) into LDS.  For a point light each sample builds its own
// cone in fp32 with the same formulas.
#include "cvr_ctx.hpp"
#include <cmath>
#include <cstring>
#include <vector>

namespace cvr {

// ----------------------------------------------------------------- cone geometry
// Per-axis literal forms of ConeZAxis / ConeYAxis / ConeXAxis (:190-430).
template <class R> struct ConeGeom {
  int axis, sa, sb;
  R interval, apos0;
  R pj[4][3];
  R scale[3];
};

// float / double overloads usable from host and device code
__host__ __device__ inline float g_abs(float x) { return fabsf(x); }
__host__ __device__ inline double g_abs(double x) { return fabs(x); }
__host__ __device__ inline float g_cos(float x) { return cosf(x); }
__host__ __device__ inline double g_cos(double x) { return cos(x); }
__host__ __device__ inline float g_sin(float x) { return sinf(x); }
__host__ __device__ inline double g_sin(double x) { return sin(x); }
__host__ __device__ inline float g_sqrt(float x) { return sqrtf(x); }
__host__ __device__ inline double g_sqrt(double x) { return sqrt(x); }
__host__ __device__ inline float g_ceil(float x) { return ceilf(x); }
__host__ __device__ inline double g_ceil(double x) { return ceil(x); }
__host__ __device__ inline float g_min(float a, float b) { return fminf(a, b); }
__host__ __device__ inline double g_min(double a, double b) { return fmin(a, b); }
__host__ __device__ inline float g_max(float a, float b) { return fmaxf(a, b); }
__host__ __device__ inline double g_max(double a, double b) { return fmax(a, b); }

template <class R> __host__ __device__ inline void nrm(R v[3]) {
  R l = g_sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= l; v[1] /= l; v[2] /= l;
}

template <class R>
__host__ __device__ inline void cone_setup(const R cv[3], R ang, R interval_vox, R initial_vox, const R scale[3],
                                           ConeGeom<R>& g) {
  R ac[3] = {g_abs(cv[0]), g_abs(cv[1]), g_abs(cv[2])};
  int axis = (ac[2] > ac[0] && ac[2] > ac[1]) ? 2 : (ac[1] > ac[0] ? 1 : 0);  // :446-453
  R pc = g_cos(ang), ps = g_sin(ang), nc = g_cos(-ang), ns = g_sin(-ang);
  g.axis = axis;
  for (int i = 0; i < 3; ++i) g.scale[i] = scale[i];
  R sgn = cv[axis] < 0 ? R(-1) : R(1);
  g.interval = interval_vox * sgn * scale[axis];
  g.apos0 = initial_vox * sgn * scale[axis];
  R(*pj)[3] = g.pj;
  if (axis == 2) {  // :198-216
    g.sa = 0; g.sb = 1;
    R py[3] = {0, cv[1], cv[2]}; nrm(py);
    R px[3] = {cv[0], 0, cv[2]}; nrm(px);
    R a0[3] = {px[0] * nc - px[2] * ns, 0, px[0] * ns + px[2] * nc};
    R a1[3] = {px[0] * pc - px[2] * ps, 0, px[0] * ps + px[2] * pc};
    R b0[3] = {0, py[1] * nc - py[2] * ns, py[1] * ns + py[2] * nc};
    R b1[3] = {0, py[1] * pc - py[2] * ps, py[1] * ps + py[2] * pc};
    nrm(a0); nrm(a1); nrm(b0); nrm(b1);
    for (int i = 0; i < 3; ++i) { pj[0][i] = a0[i]; pj[1][i] = a1[i]; pj[2][i] = b0[i]; pj[3][i] = b1[i]; }
  } else if (axis == 1) {  // :280-298
    g.sa = 0; g.sb = 2;
    R px[3] = {cv[0], cv[1], 0}; nrm(px);
    R pz[3] = {0, cv[1], cv[2]}; nrm(pz);
    R a0[3] = {px[0] * nc - px[1] * ns, px[0] * ns + px[1] * nc, 0};
    R a1[3] = {px[0] * pc - px[1] * ps, px[0] * ps + px[1] * pc, 0};
    R b0[3] = {0, pz[2] * ns + pz[1] * nc, pz[2] * nc - pz[1] * ns};
    R b1[3] = {0, pz[2] * ps + pz[1] * pc, pz[2] * pc - pz[1] * ps};
    nrm(a0); nrm(a1); nrm(b0); nrm(b1);
    for (int i = 0; i < 3; ++i) { pj[0][i] = a0[i]; pj[1][i] = a1[i]; pj[2][i] = b0[i]; pj[3][i] = b1[i]; }
  } else {  // :360-378
    g.sa = 1; g.sb = 2;
    R py[3] = {cv[0], cv[1], 0}; nrm(py);
    R pz[3] = {cv[0], 0, cv[2]}; nrm(pz);
    R a0[3] = {py[1] * ns + py[0] * nc, py[1] * nc - py[0] * ns, 0};
    R a1[3] = {py[1] * ps + py[0] * pc, py[1] * pc - py[0] * ps, 0};
    R b0[3] = {pz[2] * ns + pz[0] * nc, 0, pz[2] * nc - pz[0] * ns};
    R b1[3] = {pz[2] * ps + pz[0] * pc, 0, pz[2] * pc - pz[0] * ps};
    nrm(a0); nrm(a1); nrm(b0); nrm(b1);
    for (int i = 0; i < 3; ++i) { pj[0][i] = a0[i]; pj[1][i] = a1[i]; pj[2][i] = b0[i]; pj[3][i] = b1[i]; }
  }
}

// Box of one cone step relative to the sample position (:229-261).
template <class R>
__host__ __device__ inline void cone_box(const ConeGeom<R>& g, R apos, R o1[3], R o2[3]) {
  const int a = g.axis, sa = g.sa, sb = g.sb;
  R mean = g_abs(apos + g.interval * R(0.5));
  R pa1 = g.pj[0][sa] * (mean / g_abs(g.pj[0][a]));
  R pa2 = g.pj[1][sa] * (mean / g_abs(g.pj[1][a]));
  R pb1 = g.pj[2][sb] * (mean / g_abs(g.pj[2][a]));
  R pb2 = g.pj[3][sb] * (mean / g_abs(g.pj[3][a]));
  R a1 = g_min(pa1, pa2), a2 = g_max(pa1, pa2);
  R b1 = g_min(pb1, pb2), b2 = g_max(pb1, pb2);
  R qa = g_abs(a2 - a1) / g.scale[sa], qb = g_abs(b2 - b1) / g.scale[sb];
  R as = (g_ceil(qa) - qa) * R(0.5), bs = (g_ceil(qb) - qb) * R(0.5);
  a1 -= as * g.scale[sa]; a2 += as * g.scale[sa];
  b1 -= bs * g.scale[sb]; b2 += bs * g.scale[sb];
  o1[a] = g_min(apos, apos + g.interval); o2[a] = g_max(apos, apos + g.interval);
  o1[sa] = a1; o2[sa] = a2;
  o1[sb] = b1; o2[sb] = b2;
}

constexpr int kMaxShadowRows = 4096;

struct EbsK {
  VolGeom g;
  int w, h, d;           // SAT dims
  double sat_tex_mul[3]; // SAT texel per world unit: w / (G + 2 s)
  float scale[3], minSAT[3], maxSAT[3], minVol[3], maxVol[3];
  int apply_occ, apply_sdw, shells, shadow_type, nrows;
  float ao_radius, ka, kd, sdw_weight, sdw_angle, sdw_interval, sdw_initial, sdw_max_dist;
  float light_pos[3], light_fwd[3];
  int axis;
};

__device__ __forceinline__ double hilo(float hi, float lo) { return double(hi) + double(lo); }
// Below the first texel both taps of an axis clamp to cell 0 with the weight kept
// (CLAMP_TO_EDGE); the corner block at 0 holds cells 0 and 1, so the second is
// replaced by the first there.
__device__ __forceinline__ void sat_axis(double x, int n, int& i0, double& f, bool& below) {
  const double fl = floor(x);
  f = x - fl;
  below = fl < 0.0;
  i0 = min(max(int(fl), 0), n - 1);
}
// GL_LINEAR taps of the SAT (:74-83) are evaluated in FP64 from one of two layouts:
// layout 0, the corner blocks (one 64-byte record per tap, sat_axis); layout 1, the linear
// (hi, lo) SAT, 8 bytes per cell, where each of the four rows (y0 | y0+1, z0 | z0+1) gives
// its x-pair (x0, x0+1) in one 16-byte load.  In layout 1, x0 in [0, n-2] with the weight
// clamped to [0, 1] is CLAMP_TO_EDGE at both ends (below the first texel: cell 0; past
// the last: cell n-1); inside, the same taps and weights.
typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));
__device__ __forceinline__ void pair_axis(double x, int n, int& i0, double& f) {
  const double fl = floor(x);
  i0 = min(max(int(fl), 0), n - 2);
  f = fmin(fmax(x - double(i0), 0.0), 1.0);
}
// The texel coordinates of a box sum's corners (EvaluateSAT3D :87-100, after the border
// offset and clamp :102-110 / :181-183).
struct BoxC {
  double x1, x2, y1, y2, z1, z2;
};
__device__ __forceinline__ BoxC box_coords(const EbsK& k, f3 p1, f3 p2) {
  p1 = mk3(fminf(fmaxf(p1.x + k.scale[0], k.minSAT[0]), k.maxSAT[0]),
           fminf(fmaxf(p1.y + k.scale[1], k.minSAT[1]), k.maxSAT[1]),
           fminf(fmaxf(p1.z + k.scale[2], k.minSAT[2]), k.maxSAT[2]));
  p2 = mk3(fminf(fmaxf(p2.x + k.scale[0], k.minSAT[0]), k.maxSAT[0]),
           fminf(fmaxf(p2.y + k.scale[1], k.minSAT[1]), k.maxSAT[1]),
           fminf(fmaxf(p2.z + k.scale[2], k.minSAT[2]), k.maxSAT[2]));
  return BoxC{double(p1.x) * k.sat_tex_mul[0] - 0.5, double(p2.x) * k.sat_tex_mul[0] - 0.5,
              double(p1.y) * k.sat_tex_mul[1] - 0.5, double(p2.y) * k.sat_tex_mul[1] - 0.5,
              double(p1.z) * k.sat_tex_mul[2] - 0.5, double(p2.z) * k.sat_tex_mul[2] - 0.5};
}

// One corner tap of a box sum, loaded: its record (layout 0: the corner block, four
// float4; layout 1: its four x pairs) and FP64 weights.  tap_issue starts the loads and
// tap_eval interpolates, so that later taps' loads are in flight while earlier ones are
// interpolated.
typedef float fv4 __attribute__((ext_vector_type(4)));
template <int L> struct TapSlot {
  fv4 q[4];
  double fx, fy, fz;
  unsigned int below;  // layout 0: bit 0 (x), 1 (y), 2 (z) below the first texel
};
template <int L>
__device__ __forceinline__ void tap_issue(const void* __restrict__ sat, const EbsK& k, double x, double y, double z,
                                          TapSlot<L>& S) {
  int x0, y0, z0;
  if constexpr (L == 0) {
    bool bx, by, bz;
    sat_axis(x, k.w, x0, S.fx, bx);
    sat_axis(y, k.h, y0, S.fy, by);
    sat_axis(z, k.d, z0, S.fz, bz);
    S.below = unsigned(bx) | (unsigned(by) << 1) | (unsigned(bz) << 2);
    const fv4* r = static_cast<const fv4*>(sat) + 4 * ((size_t(z0) * k.h + y0) * k.w + x0);
#pragma unroll
    for (int i = 0; i < 4; ++i) S.q[i] = r[i];
  } else {
    pair_axis(x, k.w, x0, S.fx);
    pair_axis(y, k.h, y0, S.fy);
    pair_axis(z, k.d, z0, S.fz);
    S.below = 0u;
    const float2* sp = static_cast<const float2*>(sat);
    const size_t b = (size_t(z0) * k.h + y0) * k.w + x0, row = size_t(k.w), plane = size_t(k.w) * k.h;
    S.q[0] = *reinterpret_cast<const f4a8*>(sp + b);
    S.q[1] = *reinterpret_cast<const f4a8*>(sp + b + row);
    S.q[2] = *reinterpret_cast<const f4a8*>(sp + b + plane);
    S.q[3] = *reinterpret_cast<const f4a8*>(sp + b + plane + row);
  }
}
template <int L>
__device__ __forceinline__ double tap_eval(const TapSlot<L>& S) {
  const fv4 q0 = S.q[0], q1 = S.q[1], q2 = S.q[2], q3 = S.q[3];
  // (x, y) plane z0: q0 = (x0, y0), (x1, y0); q1 = (x0, y1), (x1, y1); plane z1: q2, q3
  double c000 = hilo(q0.x, q0.y), c100 = hilo(q0.z, q0.w), c010 = hilo(q1.x, q1.y), c110 = hilo(q1.z, q1.w);
  double c001 = hilo(q2.x, q2.y), c101 = hilo(q2.z, q2.w), c011 = hilo(q3.x, q3.y), c111 = hilo(q3.z, q3.w);
  if constexpr (L == 0) {
    // below the first texel of an axis its second cell is replaced by the first
    if (S.below & 4u) { c001 = c000; c101 = c100; c011 = c010; c111 = c110; }
    if (S.below & 2u) { c010 = c000; c110 = c100; c011 = c001; c111 = c101; }
    if (S.below & 1u) { c100 = c000; c110 = c010; c101 = c001; c111 = c011; }
  }
  const double fx = S.fx, fy = S.fy, fz = S.fz;
  const double c00 = c000 * (1 - fx) + c100 * fx, c10 = c010 * (1 - fx) + c110 * fx;
  const double c01 = c001 * (1 - fx) + c101 * fx, c11 = c011 * (1 - fx) + c111 * fx;
  const double c0 = c00 * (1 - fy) + c10 * fy, c1 = c01 * (1 - fy) + c11 * fy;
  return c0 * (1 - fz) + c1 * fz;
}
// Corner v of a box: V1..V8 of EvaluateSAT3D = (x2, y2, z2), (x1, y2, z2), (x2, y2, z1),
// (x1, y2, z1), then the same with y1; the box sum is V1 - V2 - V3 + V4 - V5 + V6 + V7 - V8.
template <int L>
__device__ __forceinline__ void corner_issue(const void* __restrict__ sat, const EbsK& k, const BoxC& c, int v,
                                             TapSlot<L>& S) {
  tap_issue<L>(sat, k, (v & 1) ? c.x1 : c.x2, (v & 4) ? c.y1 : c.y2, (v & 2) ? c.z1 : c.z2, S);
}

// A stream of box sums with four corner taps always in flight: slot j holds corner j or
// j + 4 of the current box; as each tap is interpolated its slot takes the tap four
// corners later in the stream (the next box's first four corners once the current box's
// last four are issued).  next(c, wq) gives the next box's corners and a weight and
// returns false at the end of the stream; fold(S, wq) takes the box sums in stream order,
// each summed left to right as V1 - V2 - V3 + V4 - V5 + V6 + V7 - V8.  Box geometry never
// depends on loaded values, so the next box is known before the current one is summed.
template <int L, class Next, class Fold>
__device__ __forceinline__ void box_stream(const void* __restrict__ sat, const EbsK& k, Next&& next, Fold&& fold) {
  BoxC c;
  float wq = 0.f;
  if (!next(c, wq)) return;
  TapSlot<L> s0, s1, s2, s3;
  corner_issue<L>(sat, k, c, 0, s0);
  corner_issue<L>(sat, k, c, 1, s1);
  corner_issue<L>(sat, k, c, 2, s2);
  corner_issue<L>(sat, k, c, 3, s3);
  for (;;) {
    double acc = tap_eval<L>(s0);
    corner_issue<L>(sat, k, c, 4, s0);
    acc = acc - tap_eval<L>(s1);
    corner_issue<L>(sat, k, c, 5, s1);
    acc = acc - tap_eval<L>(s2);
    corner_issue<L>(sat, k, c, 6, s2);
    acc = acc + tap_eval<L>(s3);
    corner_issue<L>(sat, k, c, 7, s3);
    float wn = 0.f;
    // the current box's corners are all issued; at the end of the stream next() leaves c
    // as it is, and the slots take the last box's corners again (in bounds, never summed),
    // so that every lane issues the same loads and no slot keeps two values
    const bool more = next(c, wn);
    acc = acc - tap_eval<L>(s0);
    corner_issue<L>(sat, k, c, 0, s0);
    acc = acc + tap_eval<L>(s1);
    corner_issue<L>(sat, k, c, 1, s1);
    acc = acc + tap_eval<L>(s2);
    corner_issue<L>(sat, k, c, 2, s2);
    acc = acc - tap_eval<L>(s3);
    corner_issue<L>(sat, k, c, 3, s3);
    fold(acc, wq);
    if (!more) break;
    wq = wn;
  }
}

// EvaluateSAT3D (:87-100) of one box.
template <int L>
__device__ __forceinline__ double box_sum(const void* __restrict__ sat, const EbsK& k, f3 p1, f3 p2) {
  bool first = true;
  double S = 0.0;
  box_stream<L>(
      sat, k,
      [&](BoxC& c, float&) {
        if (!first) return false;
        c = box_coords(k, p1, p2);
        first = false;
        return true;
      },
      [&](double v, float) { S = v; });
  return S;
}

// ExtinctionAmbientOcclusion (:112-146).
template <int L>
__device__ float ebs_ao(const void* __restrict__ sat, const EbsK& k, f3 tx) {
  const f3 sc = ld3(k.scale);
  const double rad = k.ao_radius;
  int gi = 0, fi = 0;
  double t = 0.0, Si = 0.0;
  box_stream<L>(
      sat, k,
      [&](BoxC& c, float&) {
        if (gi >= k.shells) return false;
        const float r1f = k.ao_radius * float(gi + 1);
        c = box_coords(k, tx - sc * r1f, tx + sc * r1f);
        ++gi;
        return true;
      },
      [&](double S1, float) {
        if (fi == 0) {
          t = S1 * (1.0 / (rad * rad));
        } else {
          const double r1 = rad * double(fi + 1);
          t = t + (S1 - Si) * (1.0 / (r1 * r1));
        }
        Si = S1;
        ++fi;
      });
  const double rn = rad * double(k.shells);
  return __expf(-float((1.0 / (rn * rn)) * t));
}

__device__ __forceinline__ float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// EvaluateShadowSAT3D (:148-188) summed along the cone; directional light from the table.
template <int L>
__device__ float ebs_shadow_table(const void* __restrict__ sat, const EbsK& k, const ShadowRow* rows, int nrows,
                                  f3 tx, unsigned int& nbox) {
  const float pa = comp(tx, k.axis), lo = k.minVol[k.axis], hi = k.maxVol[k.axis];
  double st = 0.0;
  int r = 0;
  box_stream<L>(
      sat, k,
      [&](BoxC& c, float& wq) {
        if (r >= nrows) return false;
        const ShadowRow& R = rows[r];
        const float nxt = pa + R.next;
        if (!(nxt > lo && nxt < hi)) return false;
        c = box_coords(k, tx + ld3(R.o1), tx + ld3(R.o2));
        wq = R.inv_vq;
        ++r;
        return true;
      },
      [&](double S, float inv_vq) {
        st += S * double(inv_vq) * double(k.sdw_weight);
        ++nbox;
      });
  return __expf(-float(st));
}

// Point light: the cone axis depends on the sample (:437-438).
template <int L>
__device__ float ebs_shadow_point(const void* __restrict__ sat, const EbsK& k, f3 tx, unsigned int& nbox) {
  f3 real = tx - mk3(0.5f * k.g.G[0], 0.5f * k.g.G[1], 0.5f * k.g.G[2]);
  f3 cvv = norm3(ld3(k.light_pos) - real);
  float cv[3] = {cvv.x, cvv.y, cvv.z};
  ConeGeom<float> cg;
  cone_setup<float>(cv, k.sdw_angle, k.sdw_interval, k.sdw_initial, k.scale, cg);
  const int a = cg.axis;
  const float pa = comp(tx, a), lo = k.minVol[a], hi = k.maxVol[a];
  float apos = cg.apos0;
  double st = 0.0;
  box_stream<L>(
      sat, k,
      [&](BoxC& c, float& vq) {
        if (!((apos / cv[a]) < k.sdw_max_dist)) return false;
        const float nxt = pa + (apos + cg.interval);
        if (!(nxt > lo && nxt < hi)) return false;
        float o1[3], o2[3];
        cone_box<float>(cg, apos, o1, o2);
        const f3 p1 = tx + ld3(o1), p2 = tx + ld3(o2);
        vq = (fabsf(p1.x - p2.x) / k.scale[0]) * (fabsf(p1.y - p2.y) / k.scale[1]) * (fabsf(p1.z - p2.z) / k.scale[2]);
        c = box_coords(k, p1, p2);
        apos = apos + cg.interval;
        return true;
      },
      [&](double S, float vq) {
        st += (S / double(vq)) * double(k.sdw_weight);
        ++nbox;
      });
  return __expf(-float(st));
}

// Directional light: the cone chain is the same for every sample.  The host builds it
// once per render in FP64 (ebs_rows: the decisions in ceil() and the distance test then
// match a double evaluation of :190-430 exactly); each workgroup copies the rows into
// LDS.  Row r is the r-th box.
__device__ __forceinline__ void ebs_load_rows(const ShadowRow* __restrict__ rows, int nrows, ShadowRow* s_rows) {
  const float4* src = reinterpret_cast<const float4*>(rows);
  float4* dst = reinterpret_cast<float4*>(s_rows);
  for (int i = threadIdx.x; i < nrows * 2; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

template <typename T, int L>
// One wave per SIMD: the box streams keep four corner records in flight per lane, which
// with the FP64 interpolation takes more than the 256 architectural VGPRs; at one wave
// per SIMD the excess lives in AGPRs instead of scratch memory (256^3 @ 1024^2 with
// shadows: 62.0 ms at two waves per SIMD with scratch spills, 53.2 ms at one; 76.2 ms
// before the streams, one box at a time at two waves per SIMD).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_ebs(RayFrame rf, EbsK k, const float4* __restrict__ tf, int ntf,
                                              const void* __restrict__ sat, const ShadowRow* __restrict__ rows,
                                              float step, float et, PhongK ph,
                                              const float4* __restrict__ grad, float4* __restrict__ out,
                                              unsigned long long* __restrict__ counters) {
  extern __shared__ float4 smem[];
  float4* s_tf = smem;
  ShadowRow* s_rows = reinterpret_cast<ShadowRow*>(smem + ntf);
  for (int i = threadIdx.x; i < ntf; i += blockDim.x) s_tf[i] = tf[i];
  ebs_load_rows(rows, k.nrows, s_rows);
  const int nrows = k.nrows;
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tiles_x = (rf.W + 15) >> 4;
  int tx_, ty_;
  super_tile(bid, tiles_x, (rf.H + 15) >> 4, 16, 32, tx_, ty_);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = tx_ * 16 + (w & 1) * 8 + (lane & 7);
  const int py = ty_ * 16 + (w >> 1) * 8 + (lane >> 3);
  if (px >= rf.W || py >= rf.H) return;

  f3 dir = camera_ray(rf, px, py);
  f3 eye = ld3(rf.eye);
  float tn, tf_;
  float4 dst = make_float4(0.f, 0.f, 0.f, 0.f);
  unsigned int nsamp = 0, nshade = 0, nbox = 0;
  const float ka = k.apply_occ ? k.ka : 0.f, kd = k.apply_sdw ? k.kd : 0.f;
  const float inv_k = 1.0f / (ka + kd);
  if (ray_box(eye, dir, k.g.G, tn, tf_)) {
    const float D = fabsf(tf_ - tn);
    const f3 tp = eye + dir * tn + mk3(0.5f * k.g.G[0], 0.5f * k.g.G[1], 0.5f * k.g.G[2]);
    for (int j = 0; j < kMaxSamples; ++j) {
      const float s = float(j) * step;
      if (!(s < D)) break;
      const float h = fminf(step, D - s);
      const f3 p = tp + dir * (s + 0.5f * h);
      const float dens = sample_volume_oct<T>(k.g, p);
      float4 src = sample_tf(s_tf, ntf, dens);
      ++nsamp;
      if (src.w > 0.0f) {
        ++nshade;
        float Io = 0.f, Is = 0.f;
        if (k.apply_occ) {
          Io = ebs_ao<L>(sat, k, p);
          nbox += unsigned(k.shells);
        }
        if (k.apply_sdw)
          Is = k.shadow_type == 1 ? ebs_shadow_table<L>(sat, k, s_rows, nrows, p, nbox)
                                  : ebs_shadow_point<L>(sat, k, p, nbox);
        if (ph.on) {
          // Phong branch (:523-543): a sample whose gradient vanishes keeps its colour
          float dd, sp;
          if (phong_terms(grad, k.g, ph, p, eye, dd, sp)) {
            const float s = Is * ((k.apply_sdw ? ph.ks : 0.f) * sp);
            src.x = inv_k * (src.x * Io * ka + Is * (src.x * kd * dd)) + s * ph.spec[0];
            src.y = inv_k * (src.y * Io * ka + Is * (src.y * kd * dd)) + s * ph.spec[1];
            src.z = inv_k * (src.z * Io * ka + Is * (src.z * kd * dd)) + s * ph.spec[2];
          }
        } else {
          // (1/(ka+kd)) * (rgb*Io*ka + rgb*Is*kd), :547
          src.x = inv_k * (src.x * Io * ka + src.x * Is * kd);
          src.y = inv_k * (src.y * Io * ka + src.y * Is * kd);
          src.z = inv_k * (src.z * Io * ka + src.z * Is * kd);
        }
        const float a = 1.0f - __expf(-src.w * h);
        const float wgt = (1.0f - dst.w) * a;
        dst.x += wgt * src.x;
        dst.y += wgt * src.y;
        dst.z += wgt * src.z;
        dst.w += wgt;
        if (dst.w > et) break;
      }
    }
  }
  out[size_t(py) * rf.W + px] = dst;
  if (counters) {
    atomicAdd(&counters[0], (unsigned long long)nsamp);
    atomicAdd(&counters[1], (unsigned long long)nshade);
    atomicAdd(&counters[3], (unsigned long long)nbox);
  }
}

// Object-space light cache (rc1pextbsd/lightcachecomputation.comp:440-468): cell i holds
// (AO, shadow) at texture-space position (i + 1/2) * cell, 1 for a term that is off,
// from the renderer's own AO and shadow evaluations.  One lane per cell.
template <int L>
// 64-lane workgroups, at most one wave per SIMD (a 32^3 cache is 512 waves): the box
// streams' registers beyond the 256 architectural VGPRs live in AGPRs.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_lc_ebs(EbsK k, const void* __restrict__ sat, const ShadowRow* __restrict__ rows,
                                                int rx, int ry, int rz, f3 cell, float2* __restrict__ lc) {
  extern __shared__ float4 smem[];
  ShadowRow* s_rows = reinterpret_cast<ShadowRow*>(smem);
  ebs_load_rows(rows, k.nrows, s_rows);
  const int nrows = k.nrows;
  const size_t n = size_t(rx) * ry * rz;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = int(i % size_t(rx)), y = int((i / size_t(rx)) % size_t(ry)), z = int(i / (size_t(rx) * ry));
  const f3 tp = mk3((float(x) + 0.5f) * cell.x, (float(y) + 0.5f) * cell.y, (float(z) + 0.5f) * cell.z);
  unsigned int nbox = 0;
  float Io = 1.f, Is = 1.f;
  if (k.apply_occ) Io = ebs_ao<L>(sat, k, tp);
  if (k.apply_sdw)
    Is = k.shadow_type == 1 ? ebs_shadow_table<L>(sat, k, s_rows, nrows, tp, nbox)
                            : ebs_shadow_point<L>(sat, k, tp, nbox);
  lc[i] = make_float2(Io, Is);
}

int launch_ebs(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st) {
  const cvr_ebs_params& P = c->ebs;
  if (P.apply_occlusion == 0 && P.apply_shadow == 0) {
    set_error("cvr_render: EBS needs occlusion or shadow enabled (the shading divides by ka+kd)");
    return -1;
  }
  if (P.ao_shells < 1 || !(P.ao_radius > 0) || !(P.sdw_interval > 0)) {
    set_error("cvr_render: EBS ao_shells >= 1, ao_radius > 0 and sdw_interval > 0 required");
    return -1;
  }
  int rc = build_sat(c, st, P.sat_layout == 0);
  if (rc) return rc;
  EbsK k{};
  k.g = geom_of(c);
  k.w = c->nx + 2; k.h = c->ny + 2; k.d = c->nz + 2;
  const double sc[3] = {double(c->sx), double(c->sy), double(c->sz)};
  const int dims[3] = {c->nx, c->ny, c->nz};
  const int sdim[3] = {k.w, k.h, k.d};
  for (int i = 0; i < 3; ++i) {
    const float s = float(sc[i]);
    const float G = k.g.G[i];
    k.scale[i] = s;
    k.minSAT[i] = s * 0.5f;
    k.maxSAT[i] = G + s * 1.5f;
    k.minVol[i] = s * 0.5f;
    k.maxVol[i] = G - s * 0.5f;
    // texel = p / (G + 2s) * (n+2) - 1/2  (GL_LINEAR over the SAT texture)
    k.sat_tex_mul[i] = double(sdim[i]) / (double(G) + 2.0 * double(s));
  }
  k.apply_occ = P.apply_occlusion; k.apply_sdw = P.apply_shadow;
  k.shells = P.ao_shells; k.shadow_type = P.shadow_type;
  k.ao_radius = P.ao_radius; k.ka = P.ka; k.kd = P.kd;
  k.sdw_weight = P.sdw_weight;
  k.sdw_angle = float(double(P.sdw_angle_deg) * 3.14159265358979323846 / 180.0);
  k.sdw_interval = P.sdw_interval; k.sdw_initial = P.sdw_initial;
  k.sdw_max_dist = P.sdw_max_dist > 0 ? P.sdw_max_dist : 0.75f * volume_diagonal(c);
  if (light) {
    for (int i = 0; i < 3; ++i) { k.light_pos[i] = light->position[i]; k.light_fwd[i] = light->forward[i]; }
  } else {
    for (int i = 0; i < 3; ++i) { k.light_pos[i] = rf.eye[i]; k.light_fwd[i] = -rf.f[i]; }
  }
  // directional light: the cone chain in FP64 on the host; |apos| grows by interval*scale
  // per row and the chain stops once |apos| / |cv_axis| >= max distance, which bounds it
  k.nrows = 0;
  if (P.apply_shadow && P.shadow_type == 1) {
    double cv[3] = {k.light_fwd[0], k.light_fwd[1], k.light_fwd[2]};
    nrm(cv);
    const double ac[3] = {std::fabs(cv[0]), std::fabs(cv[1]), std::fabs(cv[2])};
    k.axis = (ac[2] > ac[0] && ac[2] > ac[1]) ? 2 : (ac[1] > ac[0] ? 1 : 0);
    const double per_row = double(k.sdw_interval) * double(k.scale[k.axis]);
    const double n = double(k.sdw_max_dist) / per_row + std::fabs(double(k.sdw_initial)) / double(k.sdw_interval) + 2;
    if (!(n < kMaxShadowRows)) {
      set_error("cvr_render: EBS shadow chain longer than 4096 boxes");
      return -4;
    }
    const double scd[3] = {double(k.scale[0]), double(k.scale[1]), double(k.scale[2])};
    ConeGeom<double> cg;
    cone_setup<double>(cv, double(k.sdw_angle), double(k.sdw_interval), double(k.sdw_initial), scd, cg);
    std::vector<ShadowRow>& rows = c->ebs_rows;
    rows.clear();
    double apos = cg.apos0;
    for (int r = 0; r < int(n); ++r, apos = apos + cg.interval) {
      if (!((apos / cv[cg.axis]) < double(k.sdw_max_dist))) break;
      double o1[3], o2[3];
      cone_box<double>(cg, apos, o1, o2);
      ShadowRow row;
      double vq = 1.0;
      for (int i = 0; i < 3; ++i) {
        row.o1[i] = float(o1[i]);
        row.o2[i] = float(o2[i]);
        vq *= std::fabs(o1[i] - o2[i]) / scd[i];
      }
      row.next = float(apos + cg.interval);
      row.inv_vq = float(1.0 / vq);
      rows.push_back(row);
    }
    k.nrows = int(rows.size());
    if (!c->d_ebs_rows && hipMalloc(&c->d_ebs_rows, kMaxShadowRows * sizeof(ShadowRow)) != hipSuccess) {
      c->d_ebs_rows = nullptr;
      set_error("cvr_render: out of device memory (shadow rows)");
      return -5;
    }
    // the table is rewritten only when it changes: a render still in flight may read it
    if (k.nrows > 0 && (c->ebs_rows_dev.size() != rows.size() ||
                        std::memcmp(c->ebs_rows_dev.data(), rows.data(), rows.size() * sizeof(ShadowRow)) != 0)) {
      CVR_HIP_CHECK(hipStreamSynchronize(st));
      CVR_HIP_CHECK(hipMemcpy(c->d_ebs_rows, rows.data(), rows.size() * sizeof(ShadowRow), hipMemcpyHostToDevice));
      c->ebs_rows_dev = rows;
    }
  }
  const float step = P.step_size > 0 ? P.step_size : default_step(c);
  PhongK ph;
  if ((rc = phong_setup(c, P.apply_phong, P.gradient_type, P.ks, P.shininess, P.specular, rf, light, ph, st))) return rc;
  if (P.light_cache) {
    // PreComputeLightCache (ebsrenderer.cpp:441-520), then the object-space march
    if ((rc = lc_prepare(c, P.light_cache_res))) return rc;
    const int rx = c->lc_res[0], ry = c->lc_res[1], rz = c->lc_res[2];
    const int res[3] = {rx, ry, rz};
    float cs[3];
    for (int i = 0; i < 3; ++i) cs[i] = k.scale[i] * (float(dims[i]) / float(res[i]));  // VolumeScales * (dims / res)
    const f3 cell{cs[0], cs[1], cs[2]};
    const size_t n = size_t(rx) * ry * rz;
    // one wave per workgroup: a 32^3 cache is 512 waves, which then spread over every CU
    // instead of half of them (light-cache frame 0.927 -> 0.874 ms, the same cache)
    if (size_t(k.nrows) * sizeof(ShadowRow) > 64 * 1024) {
      set_error("cvr_render: EBS light cache: shadow chain of more than 2048 boxes does not fit LDS");
      return -4;
    }
    const ShadowRow* lrows = static_cast<const ShadowRow*>(c->d_ebs_rows);
    if (P.sat_layout == 1)
      hipLaunchKernelGGL(k_lc_ebs<1>, dim3(unsigned((n + 63) / 64)), dim3(64), size_t(k.nrows) * sizeof(ShadowRow), st,
                         k, static_cast<const void*>(c->d_sat), lrows, rx, ry, rz, cell, c->d_lc);
    else
      hipLaunchKernelGGL(k_lc_ebs<0>, dim3(unsigned((n + 63) / 64)), dim3(64), size_t(k.nrows) * sizeof(ShadowRow), st,
                         k, static_cast<const void*>(c->d_sat_oct), lrows, rx, ry, rz, cell, c->d_lc);
    CVR_HIP_CHECK(hipGetLastError());
    return launch_lc_render(c, rf, P.apply_occlusion, P.apply_shadow, P.ka, P.kd, step, P.et_threshold, ph, out, st);
  }
  const int tiles = ((rf.W + 15) / 16) * ((rf.H + 15) / 16);
  const size_t shm = size_t(c->tf_size) * sizeof(float4) + size_t(k.nrows) * sizeof(ShadowRow);
  if (shm > 150 * 1024) {
    set_error("cvr_render: EBS transfer function + shadow table exceed LDS");
    return -4;
  }
  unsigned long long* cnt = c->collect_stats ? c->d_counters : nullptr;
  const void* sat = P.sat_layout == 1 ? static_cast<const void*>(c->d_sat) : c->d_sat_oct;
  const ShadowRow* rows = static_cast<const ShadowRow*>(c->d_ebs_rows);
  float4* o4 = reinterpret_cast<float4*>(out);
  if (c->dtype == CVR_U8) {
    if (P.sat_layout == 1)
      hipLaunchKernelGGL((k_ebs<uint8_t, 1>), dim3(tiles), dim3(256), shm, st, rf, k, c->d_tf_rgbt, c->tf_size, sat, rows, step,
                         P.et_threshold, ph, c->d_grad, o4, cnt);
    else
      hipLaunchKernelGGL((k_ebs<uint8_t, 0>), dim3(tiles), dim3(256), shm, st, rf, k, c->d_tf_rgbt, c->tf_size, sat, rows, step,
                         P.et_threshold, ph, c->d_grad, o4, cnt);
  } else {
    if (P.sat_layout == 1)
      hipLaunchKernelGGL((k_ebs<uint16_t, 1>), dim3(tiles), dim3(256), shm, st, rf, k, c->d_tf_rgbt, c->tf_size, sat,
                         rows, step, P.et_threshold, ph, c->d_grad, o4, cnt);
    else
      hipLaunchKernelGGL((k_ebs<uint16_t, 0>), dim3(tiles), dim3(256), shm, st, rf, k, c->d_tf_rgbt, c->tf_size, sat,
                         rows, step, P.et_threshold, ph, c->d_grad, o4, cnt);
  }
  CVR_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace cvr
