// Surrogate network inference shared by the single-frame and batched renderers:
// bf16 MFMA types and layout constants, the wavefront-cooperative forward pass,
// and the per-tile surrogate march.
#pragma once
#include "cvr_cone.hpp"

namespace cvr {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4v_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8v_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2v_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef short i16x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

constexpr int kH = 64, kOut = 16;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kH * kNFeat;
constexpr int kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kH;
constexpr int kOffW4 = kOffB3 + kH, kOffB4 = kOffW4 + kOut * kH;
constexpr int kNParams = kOffB4 + kOut;  // 10448
// Packed inference weights: layer 1 as 4 A tiles of v_mfma_f32_16x16x16_bf16 (64 lanes x
// 4 bf16), layers 2, 3 and 4 as A tiles of v_mfma_f32_16x16x32_bf16 (64 lanes x 8 bf16):
// 4 x 2, 4 x 2 and 1 x 2 (row tile x 32-wide K chunk).
constexpr int kTiles1 = 4, kTiles2 = 8, kTiles3 = 8, kTiles4 = 2;
constexpr int kNTiles = kTiles1 + kTiles2 + kTiles3 + kTiles4;
constexpr int kPackShorts = kTiles1 * 64 * 4 + (kTiles2 + kTiles3 + kTiles4) * 64 * 8;

// K order of a 32-wide chunk of v_mfma_f32_16x16x32_bf16 (lane quarter q holds K = 8q..8q+7)
// mapped onto the units of the layer before, so that a layer's D tiles are the next
// layer's B operand without any lane exchange: K = 8q + e of chunk c is unit
// 16 (2c + e/4) + 4q + e%4 (rows 4q..4q+3 of D tiles 2c and 2c+1 of that lane).
__host__ __device__ constexpr int chunk_unit(int c, int q, int e) { return 16 * (2 * c + (e >> 2)) + 4 * q + (e & 3); }
constexpr int kNBias = kH + kH + kH + kOut;
constexpr int kWaveScratch = kNFeat * 64 + 2 * 64;

// fp32 -> bf16, round to nearest even (inputs are finite)
__device__ __forceinline__ short f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return short(u >> 16);
}
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(uint32_t(uint16_t(b)) << 16); }

__device__ __forceinline__ f32x4_t mfma16(bf16x4_t a, bf16x4_t b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4_t mfma32(bf16x8v_t a, bf16x8v_t b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Round two D tiles to nearest even into one 8 x bf16 B operand, a pair at a time
// (v_cvt_pk_bf16_f32), then ReLU on the bf16 bits: a negative bf16 (also -0) is a
// negative int16, so a packed 16-bit integer max with 0 (v_pk_max_i16) gives exactly
// bf16(max(x, 0)), without an fp32 max (and its input canonicalisation) per value.
__device__ __forceinline__ unsigned int relu2(float x, float y) {
  const f32x2_t p = {x, y};
  const i16x2_t v = __builtin_bit_cast(i16x2_t, __builtin_convertvector(p, bf16x2v_t));
  return __builtin_bit_cast(unsigned int, __builtin_elementwise_max(v, i16x2_t{0, 0}));
}
__device__ __forceinline__ bf16x8v_t relu_pack(f32x4_t a, f32x4_t b) {
  const u32x4_t w = {relu2(a[0], a[1]), relu2(a[2], a[3]), relu2(b[0], b[1]), relu2(b[2], b[3])};
  return __builtin_bit_cast(bf16x8v_t, w);
}

// Wavefront-cooperative inference: every lane brings one sample's 16 features; the
// 64 samples go through the network as 4 column tiles of MFMA.  sW/sB: packed weights
// and biases in LDS; sX: this wave's LDS scratch of kWaveScratch floats: the inputs as
// 64 rows of 16 bf16 (rounded to nearest even once, here), kInRow dwords apart, then
// 64x2 outputs at kNFeat * 64.  Returns outputs 0 and 1.
constexpr int kInRow = 10;  // 8 dwords of inputs + 2 of padding: the operand reads hit 32 distinct banks
static_assert(64 * kInRow <= kNFeat * 64, "input rows fit below the outputs");
__device__ __forceinline__ void mlp_wave(const short* __restrict__ sW, const float* __restrict__ sB, float* sX,
                                         const float* f, float& o0, float& o1) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
  u32x2_t* row = reinterpret_cast<u32x2_t*>(sX + lane * kInRow);
#pragma unroll
  for (int i = 0; i < kNFeat / 4; ++i) {
    const f32x2_t a = {f[4 * i], f[4 * i + 1]}, b = {f[4 * i + 2], f[4 * i + 3]};
    row[i] = u32x2_t{__builtin_bit_cast(unsigned int, __builtin_convertvector(a, bf16x2v_t)),
                     __builtin_bit_cast(unsigned int, __builtin_convertvector(b, bf16x2v_t))};
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const bf16x4_t* W1 = reinterpret_cast<const bf16x4_t*>(sW);
  const bf16x8v_t* W2 = reinterpret_cast<const bf16x8v_t*>(sW + kTiles1 * 64 * 4);
  const bf16x8v_t* W3 = W2 + kTiles2 * 64;
  const bf16x8v_t* W4 = W3 + kTiles3 * 64;
  const f32x4_t* B4 = reinterpret_cast<const f32x4_t*>(sB);  // 4 biases of rows 4q..4q+3 per (layer, row tile)
  // The four 16-sample column tiles go through the network together: every weight tile
  // is read from LDS once for all four, and the four independent MFMA chains hide each
  // other's latency.
  f32x4_t h[4][4];  // [column tile][row tile]
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    // lane 16q + r: inputs 4q..4q+3 of sample 16u + r
    const bf16x4_t xb = *reinterpret_cast<const bf16x4_t*>(sX + (u * 16 + r) * kInRow + 2 * q);
    // layer 1: K = 16
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) h[u][mt] = mfma16(W1[mt * 64 + lane], xb, B4[mt * 4 + q]);
  }
  // layers 2 and 3: K = 64 as two 32-wide chunks
#pragma unroll
  for (int l = 0; l < 2; ++l) {
    const bf16x8v_t* Wl = l == 0 ? W2 : W3;
    bf16x8v_t b0[4], b1[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      b0[u] = relu_pack(h[u][0], h[u][1]);
      b1[u] = relu_pack(h[u][2], h[u][3]);
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const bf16x8v_t wa = Wl[(mt * 2) * 64 + lane], wb = Wl[(mt * 2 + 1) * 64 + lane];
      const f32x4_t bias = B4[(l + 1) * 16 + mt * 4 + q];
#pragma unroll
      for (int u = 0; u < 4; ++u) h[u][mt] = mfma32(wb, b1[u], mfma32(wa, b0[u], bias));
    }
  }
  // layer 4: 16 outputs (rows 0, 1 used)
  const bf16x8v_t w40 = W4[lane], w41 = W4[64 + lane];
  const f32x4_t b4 = B4[48 + q];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const f32x4_t acc = mfma32(w41, relu_pack(h[u][2], h[u][3]), mfma32(w40, relu_pack(h[u][0], h[u][1]), b4));
    // lanes 0..15 (q = 0) hold outputs 0 and 1 of samples u * 16 + r
    if (q == 0) {
      sX[kNFeat * 64 + (u * 16 + r) * 2] = acc[0];
      sX[kNFeat * 64 + (u * 16 + r) * 2 + 1] = acc[1];
    }
  }
  // lanes 0..15 wrote outputs 0 and 1 of every sample; hand them back
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  o0 = sX[kNFeat * 64 + lane * 2];
  o1 = sX[kNFeat * 64 + lane * 2 + 1];
  __builtin_amdgcn_wave_barrier();
}

struct NnK {
  float lo[kNFeat], hi[kNFeat];  // training range (+- margin): outside -> exact tracer
  int exact_all;                 // validation gate failed: every sample uses the exact tracer
};

// Shape of the work of one wavefront: an (1 << tw_log2) x (64 >> tw_log2) pixel tile
// (four of them, 2 x 2, per workgroup), and how many of the wave's marching lanes
// must be waiting on a shaded sample before the wave gathers their features and runs
// the network together (`park`: a fraction of the active lanes; 0 = as soon as one
// waits).  Neither changes any pixel: every ray takes the same samples in the same
// order and the network maps each sample independently of the others.
struct NnSched {
  int tw_log2;
  float park;
  int span;  // wave tiles per wave (pixels handed out as lanes free up)
};

NnK nn_consts(cvr_ctx* c);
ConeK cone_constants(const cvr_ctx* c, const cvr_cone_params& P, const float L[3]);
void light_dir(const RayFrame& rf, const cvr_light* light, float L[3]);

// Renders one workgroup tile of frame `rf`: 2 x 2 waves, each owning `span` wave tiles
// of the NnSched shape stacked along y.  A lane whose ray ends writes its pixel and,
// once enough lanes of the wave are idle, takes the wave's next unrendered pixel, so the
// wave keeps its lanes busy until its pixels run out instead of idling behind its
// longest ray.  sW/sB/s_tf: the network and transfer function staged in LDS; sXw: this
// wave's LDS scratch.  Counters (samples, shaded, exact fallbacks) accumulate into cnt.
// kExact: a sample whose inputs leave the training range takes the exact cone tracer's
// answers; without it the tile is only marked (the return value: some lane of this wave
// met such a sample) for a second pass with kExact, and its pixels are provisional.
// Which samples a ray takes and their opacities do not depend on the shading, so both
// passes march a tile identically.
template <typename T, int F, bool kExact = true>
__device__ __forceinline__ bool nn_tile(const RayFrame& rf, const VolGeom& g, const PyrGeom& pg,
                                        const void* __restrict__ pyr, const ConeK& K, const NnK& nk,
                                        const NnSched& sc, const short* sW, const float* sB, float* sXw,
                                        const float4* s_tf, int ntf, float step, float et, float4* __restrict__ out,
                                        int tx, int ty, unsigned int* cnt) {
  const int tl = sc.tw_log2, tw = 1 << tl, th = 64 >> tl, span = sc.span;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int x0 = tx * 2 * tw + (w & 1) * tw;
  const int y0 = (ty * 2 + (w >> 1)) * th * span;
  const int total = 64 * span;
  const f3 eye = ld3(rf.eye);
  const f3 half = mk3(0.5f * g.G[0], 0.5f * g.G[1], 0.5f * g.G[2]);
  const float ka = K.apply_occ ? K.ka : 0.f, kd = K.apply_sdw ? K.kd : 0.f;
  const float inv_k = 1.0f / (ka + kd);
  const f3 L = ld3(K.L);
  const uint64_t below = (1ull << lane) - 1ull;
  // per-lane ray state (pix < 0: idle)
  int pix = -1, px = 0, py = 0, j = 0;
  float D = 0.f, h = 0.f;
  f3 dir = mk3(0.f, 0.f, 1.f), tp = eye, p = eye;  // the occlusion axis is -dir
  float4 dst = make_float4(0.f, 0.f, 0.f, 0.f), src = dst;
  // The corner block of the lane's next sample, loaded as soon as the current sample is
  // known, so that its read is in flight while the wave gathers and shades (sample j sits
  // at the midpoint of [s, s + h], s = j step, h = min(step, D - s); past the end of the
  // ray the position is still finite and its clamped block in bounds).
  typedef typename VolOct<T>::type RT;
  auto pos = [&](int jj) {
    const float s = float(jj) * step;
    return tp + dir * (s + 0.5f * fminf(step, D - s));
  };
  RT rec = {};
  bool parked = false, ood_seen = false;
  int next = 0;  // wave-uniform: next pixel of the span to hand out
  for (;;) {
    const uint64_t idle = __ballot(pix < 0);
    const int n_idle = __popcll(idle);
    if (next < total && (n_idle >= 16 || n_idle == 64)) {
      // hand the next pixels to the idle lanes, in lane order
      const int mine = next + __popcll(idle & below);
      if (pix < 0 && mine < total) {
        const int t = mine >> 6, l = mine & 63;
        px = x0 + (l & (tw - 1));
        py = y0 + t * th + (l >> tl);
        float tn = 0.f, tf_ = 0.f;
        if (px < rf.W && py < rf.H) {
          dir = camera_ray(rf, px, py);
          if (ray_box(eye, dir, g.G, tn, tf_)) {
            pix = mine;
            D = fabsf(tf_ - tn);
            tp = eye + dir * tn + half;
            j = 0;
            dst = make_float4(0.f, 0.f, 0.f, 0.f);
            rec = volume_oct_load<T>(g, pos(0));
          } else {
            out[size_t(py) * rf.W + px] = make_float4(0.f, 0.f, 0.f, 0.f);  // the ray misses the box
          }
        }
      }
      next += n_idle;
    }
    const bool busy = pix >= 0;
    if (!__any(busy)) {
      if (next >= total) break;
      continue;
    }
    // one sample of every marching lane; a lane whose sample needs shading parks on it
    // (keeps p, h, src) while the others march on
    bool finished = false;
    if (busy && !parked) {
      const float s = float(j) * step;
      if (!(s < D) || j >= kMaxSamples) {
        finished = true;
      } else {
        h = fminf(step, D - s);
        p = tp + dir * (s + 0.5f * h);
        src = sample_tf(s_tf, ntf, volume_oct_lerp<T>(g, rec, p));
        rec = volume_oct_load<T>(g, pos(j + 1));
        ++cnt[0];
        if (src.w > 0.f) parked = true;
        else ++j;
      }
    }
    // the wave shades all parked samples together once enough lanes wait, or when every
    // marching lane waits
    const int np = __popcll(__ballot(parked)), na = __popcll(__ballot(busy && !finished));
    if (np > 0 && (np == na || float(np) >= sc.park * float(na))) {
      float f[kNFeat];
      bool ood = false;
      float bo = 0.f, bs = 0.f;
      if (parked) {
        nn_feats<F>(pyr, pg, K.ft, p, mk3(-dir.x, -dir.y, -dir.z), L, f);
#pragma unroll
        for (int c = 0; c < kNFeat; ++c) ood |= (f[c] < nk.lo[c]) | (f[c] > nk.hi[c]);
        // the network-free estimate now, so the features die inside the network's input staging
        nn_base(f, float(K.n_occ - 1) * K.ring_cos, bo, bs);
      } else {
#pragma unroll
        for (int c = 0; c < kNFeat; ++c) f[c] = 0.f;
      }
      float Io, Is;
      mlp_wave(sW, sB, sXw, f, Io, Is);
      if (parked) {
        Io += bo;
        Is += bs;
        ++cnt[1];
        if (!kExact) {
          ood_seen |= ood;
        } else if (ood || nk.exact_all) {
          ++cnt[2];
          f3 k, up, right;
          occ_frame(dir, k, up, right);
          Io = K.apply_occ ? ct_occlusion<F>(pyr, pg, K, p, k, up, right) : 0.f;
          Is = K.apply_sdw ? cone_T<F>(pyr, pg, K.ss, K.n_steps, p, L) : 0.f;
        }
        Io = fminf(fmaxf(Io, 0.f), 1.f);
        Is = fminf(fmaxf(Is, 0.f), 1.f);
        const float m = inv_k * (Io * ka + Is * kd);
        const float a = 1.0f - __expf(-src.w * h);
        const float wgt = (1.0f - dst.w) * a;
        dst.x += wgt * src.x * m;
        dst.y += wgt * src.y * m;
        dst.z += wgt * src.z * m;
        dst.w += wgt;
        ++j;
        parked = false;
        if (dst.w > et) finished = true;
      }
    }
    if (finished) {
      out[size_t(py) * rf.W + px] = dst;
      pix = -1;
    }
  }
  return __any(ood_seen);
}

// Stages the packed network and the transfer function of one volume in LDS (all
// threads of the workgroup take part; the caller synchronises).
__device__ __forceinline__ void nn_stage(const short* __restrict__ wpack, const float* __restrict__ bpack,
                                         const float4* __restrict__ tf, int ntf, short* sW, float* sB, float4* s_tf) {
  for (int i = threadIdx.x; i < kPackShorts / 4; i += blockDim.x)
    reinterpret_cast<bf16x4_t*>(sW)[i] = reinterpret_cast<const bf16x4_t*>(wpack)[i];
  for (int i = threadIdx.x; i < kNBias; i += blockDim.x) sB[i] = bpack[i];
  for (int i = threadIdx.x; i < ntf; i += blockDim.x) s_tf[i] = tf[i];
}

}  // namespace cvr
