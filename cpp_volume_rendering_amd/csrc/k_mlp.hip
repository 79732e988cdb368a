// Learned occlusion/shadow surrogate: bf16 MFMA training and inference, gfx950.
//
// Network: 16 -> 64 -> 64 -> 64 -> 16 (outputs 0, 1 = ambient occlusion, shadow
// transmittance), ReLU on the hidden layers, bf16 operands, fp32 accumulation.
// Parameters (fp32 master copy) are flattened as W1[64][16], b1[64], W2[64][64], b2,
// W3[64][64], b3, W4[16][64], b4[16], row = output feature.
//
// MFMA: v_mfma_f32_16x16x16_bf16.  Lane l (r = l%16, q = l/16) holds A[r][4q+e],
// B[4q+e][r] and D[4q+e][r], e = 0..3.  Activations are kept feature-major
// (features x samples), so a layer's D tile is, after bias + ReLU + bf16 rounding,
// exactly the B operand of the next layer: inference keeps all activations in
// registers; only the input features pass through LDS once.
//
// Training (k_mlp_train, k_mlp_update): a workgroup of kTrainWaves = 4 wavefronts takes a
// 64-sample chunk, 16 samples per wavefront (8 wavefronts measured 13.1 us per step
// against 12.1 us).  Each wavefront runs the forward on the register path
// and the activation gradients through W^T tiles in registers; the weight gradients are
// MFMA contractions over the chunk's 64 samples from sample-major bf16 copies in LDS that
// all four wavefronts fill, each wavefront computing a quarter of the output tiles, so the
// workgroup writes one gradient slab (no atomics).  k_mlp_update sums the slabs, takes the
// Adam step and writes the bf16 packs of the new weights.
#include "cvr_nn.hpp"
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace cvr {

// Tile (m0.., n0..) += sum_k A(m,k) B(k,n) with A(m,k) = a[m*am + k*ak], B(k,n) = b[k*bk + n*bn].
__device__ __forceinline__ f32x4_t tile_gemm(const float* a, int am, int ak, const float* b, int bk, int bn, int m0,
                                             int n0, int K, f32x4_t acc) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4_t av, bv;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = k0 + 4 * q + e;
      av[e] = f2bf(a[(m0 + r) * am + k * ak]);
      bv[e] = f2bf(b[k * bk + (n0 + r) * bn]);
    }
    acc = mfma16(av, bv, acc);
  }
  return acc;
}

// Forward of one layer for 64 samples held feature-major in LDS.
__device__ void layer_fwd(const float* __restrict__ W, const float* __restrict__ bias, int in, int out,
                          const float* aIn, float* aOut, bool relu) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  for (int mt = 0; mt < out / 16; ++mt)
    for (int nt = 0; nt < 4; ++nt) {
      f32x4_t acc;
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[e] = bias[mt * 16 + 4 * q + e];
      acc = tile_gemm(W, in, 1, aIn, 64, 1, mt * 16, nt * 16, in, acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float v = acc[e];
        aOut[(mt * 16 + 4 * q + e) * 64 + nt * 16 + r] = relu ? fmaxf(v, 0.f) : v;
      }
    }
}

// ---------------------------------------------------------------- training (MFMA register path)
// One bf16 copy of the network for training, in MFMA A-operand order:
//   [forward tiles, kPackShorts, as mlp_wave reads them] [fp32 biases, kNBias]
//   [backward tiles of W^T, kBwdShorts]
// Backward tiles, the A operands of the activation gradients G_in = W^T G_out: layer 4,
// W4^T (64 inputs x K = 16 outputs), 4 tiles of 16x16x16 where lane (r, q) holds
// W4[4q + e][16 mt + r]; layers 3 and 2, W^T (64 x K = 64), 4 x 2 tiles of 16x16x32 where
// lane (r, q) holds W[chunk_unit(c, q, e)][16 mt + r] -- the K order in which a layer's
// gradient D tiles pack into the next B operand, as the forward's activations do.  Layer
// 1's input gradient is never needed.  The update kernel writes these packs after every
// optimiser step, so the training kernel only copies them into LDS.
constexpr int kBTiles4 = 4, kBTiles3 = 8, kBTiles2 = 8;
constexpr int kBwdShorts = kBTiles4 * 64 * 4 + (kBTiles3 + kBTiles2) * 64 * 8;
constexpr int kBwdBase3 = kBTiles4 * 64 * 4, kBwdBase2 = kBwdBase3 + kBTiles3 * 64 * 8;
constexpr size_t kTrainPackBytes = size_t(kPackShorts) * 2 + size_t(kNBias) * 4 + size_t(kBwdShorts) * 2;
// Per-chunk gradient slab: the kNParams gradients of a chunk's samples, then their
// squared-error sum.  A chunk is one training workgroup's samples (four wavefronts of 16):
// a 16384-sample step is 256 workgroups, one per CU, one wavefront per SIMD.
constexpr int kSlab = kNParams + 16;
constexpr int kTrainWaves = 4;
constexpr int kTrainChunk = 16 * kTrainWaves;
// LDS row stride (bf16) of the sample-major copies of activations and gradients: the
// chunk's samples + 8 of padding, so that 16 rows read at one column fall on distinct banks.
constexpr int kTS = kTrainChunk + 8;

// Pack slots of flattened parameter i: short index in the forward pack (or -1), in the
// backward pack (or -1), float index in the bias pack (or -1).
__host__ __device__ inline void pack_slots(int i, int& fwd, int& bwd, int& bias) {
  fwd = bwd = bias = -1;
  if (i < kOffB1) {  // W1[o][k], 64 x 16: 16x16x16 tiles, lane (r, q) holds W1[16 t + r][4q + e]
    const int o = i / kNFeat, k = i % kNFeat;
    fwd = ((o / 16) * 64 + 16 * (k / 4) + o % 16) * 4 + k % 4;
    return;
  }
  if (i < kOffW2) { bias = i - kOffB1; return; }
  int l, off;
  if (i < kOffB2) { l = 2; off = kOffW2; }
  else if (i < kOffW3) { bias = kH + i - kOffB2; return; }
  else if (i < kOffB3) { l = 3; off = kOffW3; }
  else if (i < kOffW4) { bias = 2 * kH + i - kOffB3; return; }
  else if (i < kOffB4) { l = 4; off = kOffW4; }
  else { bias = 3 * kH + i - kOffB4; return; }
  const int j = i - off, o = j / kH, k = j % kH;  // W_l[o][k] over 64 inputs
  // chunk_unit(c, q, e) = k: c = k / 32, q = (k % 16) / 4, e = 4 ((k % 32) / 16) + k % 4
  const int c = k / 32, q = (k % 16) / 4, e = 4 * ((k % 32) / 16) + k % 4;
  const int t = (l == 2 ? 0 : l == 3 ? kTiles2 : kTiles2 + kTiles3) + (o / 16) * 2 + c;
  fwd = kTiles1 * 64 * 4 + (t * 64 + 16 * q + o % 16) * 8 + e;
  if (l == 4) {  // W4^T: row = input k, K = output o (16x16x16)
    bwd = ((k / 16) * 64 + 16 * (o / 4) + k % 16) * 4 + o % 4;
  } else {       // W^T: row = input k, K = output o in chunk order (16x16x32)
    const int c2 = o / 32, q2 = (o % 16) / 4, e2 = 4 * ((o % 32) / 16) + o % 4;
    bwd = (l == 3 ? kBwdBase3 : kBwdBase2) + (((k / 16) * 2 + c2) * 64 + 16 * q2 + k % 16) * 8 + e2;
  }
}

__device__ __forceinline__ short bf16_bits(float x) {
  return __builtin_bit_cast(short, __builtin_convertvector(f32x2_t{x, 0.f}, bf16x2v_t)[0]);
}
// fp32 D tiles a (rows 4q..4q+3 of tile 2c) and b (of tile 2c+1) as one 8 x bf16 B operand
// (hardware round to nearest even), the order relu_pack uses, without the ReLU.
__device__ __forceinline__ bf16x8v_t grad_pack(f32x4_t a, f32x4_t b) {
  const f32x2_t p0 = {a[0], a[1]}, p1 = {a[2], a[3]}, p2 = {b[0], b[1]}, p3 = {b[2], b[3]};
  const u32x4_t w = {__builtin_bit_cast(unsigned int, __builtin_convertvector(p0, bf16x2v_t)),
                     __builtin_bit_cast(unsigned int, __builtin_convertvector(p1, bf16x2v_t)),
                     __builtin_bit_cast(unsigned int, __builtin_convertvector(p2, bf16x2v_t)),
                     __builtin_bit_cast(unsigned int, __builtin_convertvector(p3, bf16x2v_t))};
  return __builtin_bit_cast(bf16x8v_t, w);
}
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// D tile d of column tile u (rows unit0 + 4q + e, column = sample 16u + r) into the
// sample-major LDS copy sT[unit][sample], masked by the post-ReLU activation bits `act`
// (element e of the tile's rows; nullptr: no mask).  Returns the masked tile.
__device__ __forceinline__ f32x4_t mask_tile(f32x4_t d, const short* act4) {
#pragma unroll
  for (int e = 0; e < 4; ++e) d[e] = act4[e] > 0 ? d[e] : 0.f;
  return d;
}
__device__ __forceinline__ void store_tile(short* sT, f32x4_t d, int unit0, int u) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int e = 0; e < 4; ++e) sT[(unit0 + 4 * q + e) * kTS + 16 * u + r] = bf16_bits(d[e]);
}
// An activation B operand (8 x bf16 of chunk c of column tile u) into sT[unit][sample].
__device__ __forceinline__ void store_act(short* sT, bf16x8v_t a, int c, int u) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  const i16x2_t* v = reinterpret_cast<const i16x2_t*>(&a);
#pragma unroll
  for (int e = 0; e < 8; ++e) sT[chunk_unit(c, q, e) * kTS + 16 * u + r] = v[e >> 1][e & 1];
}

// Weight and bias gradients of one layer over a chunk's samples: dW[o][k] = sum_s
// G[o][s] H[k][s] and db[o] = sum_s G[o][s] from the sample-major LDS copies sG (`out`
// rows) and sH (`in` rows), as 16x16x32 MFMA over 32-sample steps; written to the slab
// (dW row-major at offW, db at offB).  The layer's 16x16 tiles t = mt + (out / 16) nt are
// dealt to the wavefronts round robin (wavefront w: t = w, w + kTrainWaves, ...), so a
// wavefront keeps one row tile of a 64-row layer; the bias of row tile mt goes with its
// tile nt = 0.  The slab stores are non-temporal: k_mlp_update reads the slabs from other
// XCDs, whose L2s do not see this one's, so cached stores only had to be written back at
// the end of the kernel (12.2 -> 9.5 us per 16384-sample step).
__device__ __forceinline__ void weight_grads(const short* sG, const short* sH, int out, int in, float* slab, int offW,
                                             int offB, int w) {
  constexpr int kK = kTrainChunk / 32;
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  typedef short s16x8_t __attribute__((ext_vector_type(8)));
  const s16x8_t ones_bits = {0x3F80, 0x3F80, 0x3F80, 0x3F80, 0x3F80, 0x3F80, 0x3F80, 0x3F80};  // bf16 1.0
  const bf16x8v_t ones = __builtin_bit_cast(bf16x8v_t, ones_bits);
  const int mts = out / 16, ntiles = mts * (in / 16);
  for (int t = w; t < ntiles; t += kTrainWaves) {
    const int mt = t % mts, nt = t / mts;
    bf16x8v_t ga[kK];
#pragma unroll
    for (int kc = 0; kc < kK; ++kc)
      ga[kc] = *reinterpret_cast<const bf16x8v_t*>(sG + (16 * mt + r) * kTS + 32 * kc + 8 * q);
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < kK; ++kc)
      acc = mfma32(ga[kc], *reinterpret_cast<const bf16x8v_t*>(sH + (16 * nt + r) * kTS + 32 * kc + 8 * q), acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) __builtin_nontemporal_store(acc[e], &slab[offW + (16 * mt + 4 * q + e) * in + 16 * nt + r]);
    if (nt == 0) {
      f32x4_t db = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kc = 0; kc < kK; ++kc) db = mfma32(ga[kc], ones, db);
      if (r == 0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) __builtin_nontemporal_store(db[e], &slab[offB + 16 * mt + 4 * q + e]);
      }
    }
  }
}

// One optimiser step's gradients: a workgroup of kTrainWaves wavefronts takes a chunk of
// kTrainChunk samples, 16 per wavefront (one MFMA column tile each).
// X: [n][16], Tg: [n][2]; slab c (kSlab floats) receives d/dparams of
// sum_s (y_s - t_s)^2 / (2 norm_n) over the chunk's samples, and that sum.  Forward and
// activation gradients stay in registers (the renderer's MFMA register path, then W^T
// tiles in the same K order); the weight-gradient contractions over the chunk read
// sample-major bf16 copies in LDS (wavefront w's samples in columns 16w..16w+15).
__global__ __launch_bounds__(64 * kTrainWaves) void k_mlp_train(const unsigned char* __restrict__ pack,
                                                                const float* __restrict__ X,
                                                                const float* __restrict__ Tg, int n, float norm_n,
                                                                float* __restrict__ slabs) {
  __shared__ __align__(16) float sXall[kTrainWaves][16 * kInRow];
  __shared__ __align__(16) short sG[kH * kTS];
  __shared__ __align__(16) short sH[kH * kTS];
  __shared__ float sLoss[kTrainWaves];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, r = lane & 15, q = lane >> 4;
  // Every lane's weight operands (forward tiles, W^T tiles, its biases) straight from the
  // packs into registers, all loads in flight at once: with one wavefront per SIMD the
  // registers are there, and no LDS staging or barrier precedes the forward.
  const bf16x4_t* W1g = reinterpret_cast<const bf16x4_t*>(pack);
  const bf16x8v_t* W2g = reinterpret_cast<const bf16x8v_t*>(pack + size_t(kTiles1) * 64 * 4 * 2);
  const bf16x8v_t* W3g = W2g + kTiles2 * 64;
  const bf16x8v_t* W4g = W3g + kTiles3 * 64;
  const f32x4_t* B4g = reinterpret_cast<const f32x4_t*>(pack + size_t(kPackShorts) * 2);
  const unsigned char* bwd = pack + size_t(kPackShorts) * 2 + size_t(kNBias) * 4;
  const bf16x4_t* V4g = reinterpret_cast<const bf16x4_t*>(bwd);
  const bf16x8v_t* V3g = reinterpret_cast<const bf16x8v_t*>(bwd + size_t(kBwdBase3) * 2);
  const bf16x8v_t* V2g = reinterpret_cast<const bf16x8v_t*>(bwd + size_t(kBwdBase2) * 2);
  bf16x4_t w1[4], v4[4];
  bf16x8v_t w2[8], w3[8], w4[2], v3[8], v2[8];
  f32x4_t b1[4], b2[4], b3[4], b4;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    w1[t] = W1g[t * 64 + lane];
    v4[t] = V4g[t * 64 + lane];
    b1[t] = B4g[t * 4 + q];
    b2[t] = B4g[16 + t * 4 + q];
    b3[t] = B4g[32 + t * 4 + q];
  }
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    w2[t] = W2g[t * 64 + lane];
    w3[t] = W3g[t * 64 + lane];
    v3[t] = V3g[t * 64 + lane];
    v2[t] = V2g[t * 64 + lane];
  }
  w4[0] = W4g[lane];
  w4[1] = W4g[64 + lane];
  b4 = B4g[48 + q];
  const int s0 = blockIdx.x * kTrainChunk + 16 * w;  // this wavefront's first sample
  const int s = s0 + lane;
  const bool valid = lane < 16 && s < n;
  float f[kNFeat];
  if (valid) {
    // a sample's 16 inputs are 64 contiguous bytes: four 16-byte loads
    const float4* xr = reinterpret_cast<const float4*>(X + size_t(s) * kNFeat);
#pragma unroll
    for (int i = 0; i < kNFeat / 4; ++i) {
      const float4 v = xr[i];
      f[4 * i] = v.x;
      f[4 * i + 1] = v.y;
      f[4 * i + 2] = v.z;
      f[4 * i + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < kNFeat; ++i) f[i] = 0.f;
  }
  // the targets of the samples whose output rows this lane holds (lanes q = 0), read now
  // so that their latency does not sit between the forward and the backward
  const int su = s0 + r;
  const bool has_t = q == 0 && su < n;
  float t0 = 0.f, t1 = 0.f;
  if (has_t) {
    t0 = Tg[size_t(su) * 2];
    t1 = Tg[size_t(su) * 2 + 1];
  }
  // inputs: bf16 rows for the forward operand (as mlp_wave)
  float* sX = sXall[w];
  typedef unsigned int u32x2v_t __attribute__((ext_vector_type(2)));
  if (lane < 16) {
    u32x2v_t* row = reinterpret_cast<u32x2v_t*>(sX + lane * kInRow);
#pragma unroll
    for (int i = 0; i < kNFeat / 4; ++i) {
      const f32x2_t a = {f[4 * i], f[4 * i + 1]}, b = {f[4 * i + 2], f[4 * i + 3]};
      row[i] = u32x2v_t{__builtin_bit_cast(unsigned int, __builtin_convertvector(a, bf16x2v_t)),
                        __builtin_bit_cast(unsigned int, __builtin_convertvector(b, bf16x2v_t))};
    }
  }
  wave_sync();  // this wavefront's input rows in LDS
  // ---- forward, activations kept as bf16 B operands
  f32x4_t h[4];
  bf16x8v_t a1[2], a2[2], a3[2];
  {
    const bf16x4_t xb = *reinterpret_cast<const bf16x4_t*>(sX + r * kInRow + 2 * q);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) h[mt] = mfma16(w1[mt], xb, b1[mt]);
    a1[0] = relu_pack(h[0], h[1]);
    a1[1] = relu_pack(h[2], h[3]);
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
    h[mt] = mfma32(w2[mt * 2 + 1], a1[1], mfma32(w2[mt * 2], a1[0], b2[mt]));
  a2[0] = relu_pack(h[0], h[1]);
  a2[1] = relu_pack(h[2], h[3]);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
    h[mt] = mfma32(w3[mt * 2 + 1], a2[1], mfma32(w3[mt * 2], a2[0], b3[mt]));
  a3[0] = relu_pack(h[0], h[1]);
  a3[1] = relu_pack(h[2], h[3]);
  // ---- output gradient g4 = (y - t) / norm_n on outputs 0, 1 (rows e < 2 of lanes q = 0)
  float* slab = slabs + size_t(blockIdx.x) * kSlab;
  const float inv_n = 1.0f / norm_n;
  float lsum = 0.f;
  f32x4_t g4 = {0.f, 0.f, 0.f, 0.f};
  {
    const f32x4_t y = mfma32(w4[1], a3[1], mfma32(w4[0], a3[0], b4));
    if (has_t) {
      const float e0 = y[0] - t0, e1 = y[1] - t1;
      lsum = e0 * e0 + e1 * e1;
      g4[0] = e0 * inv_n;
      g4[1] = e1 * inv_n;
    }
  }
  // ---- layer 4: dW4 = G4 H3^T, then G3 = (W4^T G4) masked by h3 > 0
  store_tile(sG, g4, 0, w);
  store_act(sH, a3[0], 0, w);
  store_act(sH, a3[1], 1, w);
  __syncthreads();
  weight_grads(sG, sH, kOut, kH, slab, kOffW4, kOffB4, w);
  f32x4_t g[4];
  {
    const f32x2_t p0 = {g4[0], g4[1]}, p1 = {g4[2], g4[3]};
    typedef unsigned int u32x2_t2 __attribute__((ext_vector_type(2)));
    const u32x2_t2 gb = {__builtin_bit_cast(unsigned int, __builtin_convertvector(p0, bf16x2v_t)),
                         __builtin_bit_cast(unsigned int, __builtin_convertvector(p1, bf16x2v_t))};
    const bf16x4_t gop = __builtin_bit_cast(bf16x4_t, gb);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const f32x4_t z = {0.f, 0.f, 0.f, 0.f};
      const short* act = reinterpret_cast<const short*>(&a3[mt >> 1]) + 4 * (mt & 1);
      g[mt] = mask_tile(mfma16(v4[mt], gop, z), act);
    }
  }
  // ---- layers 3 and 2: dW = G H^T, then G_in = (W^T G) masked by the layer below
#pragma unroll
  for (int l = 3; l >= 2; --l) {
    __syncthreads();  // every wavefront's reads of sG / sH for the layer above are done
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) store_tile(sG, g[mt], 16 * mt, w);
    const bf16x8v_t* below = l == 3 ? a2 : a1;
    store_act(sH, below[0], 0, w);
    store_act(sH, below[1], 1, w);
    __syncthreads();
    weight_grads(sG, sH, kH, kH, slab, l == 3 ? kOffW3 : kOffW2, l == 3 ? kOffB3 : kOffB2, w);
    const bf16x8v_t* V = l == 3 ? v3 : v2;
    const bf16x8v_t gb0 = grad_pack(g[0], g[1]), gb1 = grad_pack(g[2], g[3]);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const f32x4_t z = {0.f, 0.f, 0.f, 0.f};
      g[mt] = mask_tile(mfma32(V[mt * 2 + 1], gb1, mfma32(V[mt * 2], gb0, z)),
                        reinterpret_cast<const short*>(&below[mt >> 1]) + 4 * (mt & 1));
    }
  }
  // ---- layer 1: dW1 = G1 X^T
  __syncthreads();
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) store_tile(sG, g[mt], 16 * mt, w);
  if (lane < 16) {
#pragma unroll
    for (int i = 0; i < kNFeat; ++i) sH[i * kTS + 16 * w + lane] = bf16_bits(f[i]);
  }
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off, 64);
  if (lane == 0) sLoss[w] = lsum;
  __syncthreads();
  weight_grads(sG, sH, kH, kNFeat, slab, kOffW1, kOffB1, w);
  if (tid == 0) {
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < kTrainWaves; ++k) t += sLoss[k];
    slab[kNParams] = t * 0.5f * inv_n;
  }
}

// Sum of the gradient slabs; mode 0 writes it (and the loss) to grad / loss, mode 1 also
// takes an Adam step (Kingma & Ba 2015) on the fp32 master parameters, mode 2 only
// packs; modes 1 and 2 write the bf16 training packs and the renderer's pack (forward
// tiles + fp32 biases) of the updated parameters.  A workgroup takes 32 parameters
// (one cache line of each slab) with 32 threads per parameter, each summing every 32nd
// slab with up to 8 loads in flight (one batch at the default 256 slabs: the kernel is
// latency-bound, 4.4 us against 6.7 us with 8 threads per parameter); the partial sums
// meet in LDS.
constexpr int kUpdParams = 32, kUpdParts = 32;
__global__ __launch_bounds__(kUpdParams * kUpdParts) void k_mlp_update(const float* __restrict__ slabs, int nslabs, int mode,
                                                    float* __restrict__ prm, float* __restrict__ grad,
                                                    float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                    float b2, float eps, float bc1, float bc2,
                                                    unsigned char* __restrict__ tpack, short* __restrict__ rpack,
                                                    float* __restrict__ loss) {
  __shared__ float part[kUpdParts][kUpdParams];
  const int pi = threadIdx.x % kUpdParams, pt = threadIdx.x / kUpdParams;
  const int i = blockIdx.x * kUpdParams + pi;  // parameter (kNParams: the loss)
  float gs = 0.f;
  if (mode != 2 && i <= kNParams) {
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int c = pt;
    for (; c + 7 * kUpdParts < nslabs; c += 8 * kUpdParts) {
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] += slabs[size_t(c + k * kUpdParts) * kSlab + i];
    }
    for (; c < nslabs; c += kUpdParts) acc[0] += slabs[size_t(c) * kSlab + i];
    gs = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  }
  part[pt][pi] = gs;
  __syncthreads();
  if (pt != 0 || i > kNParams) return;
#pragma unroll
  for (int k = 1; k < kUpdParts; ++k) gs += part[k][pi];
  if (i == kNParams) {
    if (mode != 2) *loss = gs;
    return;
  }
  float w = prm[i];
  if (mode != 2) {
    if (mode == 0) {
      grad[i] = gs;
      return;
    }
    const float mi = b1 * m[i] + (1.f - b1) * gs;
    const float vi = b2 * v[i] + (1.f - b2) * gs * gs;
    m[i] = mi;
    v[i] = vi;
    w -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    prm[i] = w;
  }
  int fwd, bwd, bias;
  pack_slots(i, fwd, bwd, bias);
  short* tf = reinterpret_cast<short*>(tpack);
  float* tb = reinterpret_cast<float*>(tpack + size_t(kPackShorts) * 2);
  short* tw = reinterpret_cast<short*>(tpack + size_t(kPackShorts) * 2 + size_t(kNBias) * 4);
  float* rb = reinterpret_cast<float*>(rpack + kPackShorts);
  if (fwd >= 0) {
    const short b = f2bf(w);
    tf[fwd] = b;
    rpack[fwd] = b;
  }
  if (bwd >= 0) tw[bwd] = f2bf(w);
  if (bias >= 0) {
    tb[bias] = w;
    rb[bias] = w;
  }
}

__global__ __launch_bounds__(64) void k_mlp_eval(const float* __restrict__ prm, const float* __restrict__ X, int n,
                                                 float* __restrict__ Y) {
  __shared__ float sX[16 * 64], sH1[64 * 64], sH2[64 * 64], sY[16 * 64];
  const int lane = threadIdx.x;
  const int s = blockIdx.x * 64 + lane;
  const bool valid = s < n;
  for (int f = 0; f < kNFeat; ++f) sX[f * 64 + lane] = valid ? X[size_t(s) * kNFeat + f] : 0.f;
  __syncthreads();
  layer_fwd(prm + kOffW1, prm + kOffB1, kNFeat, kH, sX, sH1, true);
  __syncthreads();
  layer_fwd(prm + kOffW2, prm + kOffB2, kH, kH, sH1, sH2, true);
  __syncthreads();
  layer_fwd(prm + kOffW3, prm + kOffB3, kH, kH, sH2, sH1, true);
  __syncthreads();
  layer_fwd(prm + kOffW4, prm + kOffB4, kH, kOut, sH1, sY, false);
  __syncthreads();
  if (valid) {
    Y[size_t(s) * 2] = sY[lane];
    Y[size_t(s) * 2 + 1] = sY[64 + lane];
  }
}

// Packs the fp32 master weights into the bf16 A-operand tiles mlp_wave reads (layer 1:
// 4 consecutive k per lane; layers 2-4: the 8 units chunk_unit gives per lane) and the
// biases into fp32.
__global__ void k_mlp_pack(const float* __restrict__ prm, short* __restrict__ wpack, float* __restrict__ bpack) {
  const int t = blockIdx.x;  // tile
  const int lane = threadIdx.x, r = lane & 15, q = lane >> 4;
  if (t < kTiles1) {
    for (int e = 0; e < 4; ++e) wpack[(t * 64 + lane) * 4 + e] = f2bf(prm[kOffW1 + (t * 16 + r) * kNFeat + 4 * q + e]);
  } else {
    int u = t - kTiles1, off, mt, c;
    if (u < kTiles2) { off = kOffW2; mt = u / 2; c = u % 2; }
    else if (u < kTiles2 + kTiles3) { u -= kTiles2; off = kOffW3; mt = u / 2; c = u % 2; }
    else { u -= kTiles2 + kTiles3; off = kOffW4; mt = 0; c = u; }
    short* dst = wpack + kTiles1 * 64 * 4 + (size_t(t - kTiles1) * 64 + lane) * 8;
    for (int e = 0; e < 8; ++e) dst[e] = f2bf(prm[off + (mt * 16 + r) * kH + chunk_unit(c, q, e)]);
  }
  if (t == 0) {
    for (int i = lane; i < kNBias; i += 64) {
      const int o = i < 3 * kH ? (i < kH ? kOffB1 + i : (i < 2 * kH ? kOffB2 + i - kH : kOffB3 + i - 2 * kH)) : kOffB4 + i - 3 * kH;
      bpack[i] = prm[o];
    }
  }
}

// Batch inference through the register path (the one the renderer uses).
__global__ __launch_bounds__(256) void k_mlp_infer(const short* __restrict__ wpack, const float* __restrict__ bpack,
                                                   const float* __restrict__ X, int n, float* __restrict__ Y) {
  __shared__ __align__(16) short sW[kPackShorts];
  __shared__ __align__(16) float sB[kNBias];
  __shared__ __align__(16) float sX[4][kWaveScratch];
  for (int i = threadIdx.x; i < kPackShorts / 4; i += blockDim.x)
    reinterpret_cast<bf16x4_t*>(sW)[i] = reinterpret_cast<const bf16x4_t*>(wpack)[i];
  for (int i = threadIdx.x; i < kNBias; i += blockDim.x) sB[i] = bpack[i];
  __syncthreads();
  const int w = threadIdx.x >> 6;
  const int s = blockIdx.x * 256 + threadIdx.x;
  float f[kNFeat];
  for (int i = 0; i < kNFeat; ++i) f[i] = s < n ? X[size_t(s) * kNFeat + i] : 0.f;
  float o0, o1;
  mlp_wave(sW, sB, sX[w], f, o0, o1);
  if (s < n) {
    Y[size_t(s) * 2] = o0;
    Y[size_t(s) * 2 + 1] = o1;
  }
}

// ---------------------------------------------------------------- training samples
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ float u01(uint64_t& s) {
  s += 0x9E3779B97F4A7C15ull;
  return float(mix64(s) >> 40) * (1.0f / 16777216.0f);
}
__device__ __forceinline__ f3 unit_sphere(uint64_t& s) {
  const float z = 2.f * u01(s) - 1.f, ph = 6.283185307179586f * u01(s);
  const float r = sqrtf(fmaxf(0.f, 1.f - z * z));
  return mk3(r * cosf(ph), r * sinf(ph), z);
}

// Sample i: a position drawn uniformly in the volume box until the transfer function
// gives it nonzero extinction (the samples the renderer shades), a view direction and
// a light direction uniform on the sphere; outputs the 16 features and the exact
// (occlusion, shadow) of the cone tracer minus the network-free estimate.
template <typename T, int F>
__global__ __launch_bounds__(256) void k_nn_samples(VolGeom g, PyrGeom pg, const void* __restrict__ pyr, ConeK K,
                                                    const float4* __restrict__ tf, int ntf, uint64_t seed, int n,
                                                    float* __restrict__ X, float* __restrict__ Tg,
                                                    float* __restrict__ pos_out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uint64_t s = mix64(seed ^ (uint64_t(i) * 0xD1B54A32D192ED03ull));
  f3 p = mk3(0.5f * g.G[0], 0.5f * g.G[1], 0.5f * g.G[2]);
  for (int t = 0; t < 64; ++t) {
    const f3 c = mk3(u01(s) * g.G[0], u01(s) * g.G[1], u01(s) * g.G[2]);
    p = c;
    if (sample_tf(tf, ntf, sample_volume_oct<T>(g, c)).w > 0.f) break;
  }
  const f3 dir = unit_sphere(s);
  const f3 L = unit_sphere(s);
  f3 k, up, right;
  occ_frame(dir, k, up, right);
  float f[kNFeat];
  nn_feats<F>(pyr, pg, K.ft, p, k, L, f);
  for (int c = 0; c < kNFeat; ++c) X[size_t(i) * kNFeat + c] = f[c];
  float bo, bs;
  nn_base(f, float(K.n_occ - 1) * K.ring_cos, bo, bs);
  // targets are the residuals of the exact tracer to the network-free estimate
  Tg[size_t(i) * 2] = ct_occlusion<F>(pyr, pg, K, p, k, up, right) - bo;
  Tg[size_t(i) * 2 + 1] = cone_T<F>(pyr, pg, K.ss, K.n_steps, p, L) - bs;
  if (pos_out) {
    pos_out[size_t(i) * 6 + 0] = p.x; pos_out[size_t(i) * 6 + 1] = p.y; pos_out[size_t(i) * 6 + 2] = p.z;
    pos_out[size_t(i) * 6 + 3] = dir.x; pos_out[size_t(i) * 6 + 4] = dir.y; pos_out[size_t(i) * 6 + 5] = dir.z;
  }
}

// Per-feature range of the training set (features are >= 0, so the float bit pattern
// orders like the value).
__global__ void k_feat_range(const float* __restrict__ X, int n, unsigned int* __restrict__ lo,
                             unsigned int* __restrict__ hi) {
  __shared__ unsigned int slo[kNFeat], shi[kNFeat];
  if (threadIdx.x < kNFeat) { slo[threadIdx.x] = 0x7F800000u; shi[threadIdx.x] = 0u; }
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int c = 0; c < kNFeat; ++c) {
      const unsigned int b = __float_as_uint(fmaxf(X[size_t(i) * kNFeat + c], 0.f));
      atomicMin(&slo[c], b);
      atomicMax(&shi[c], b);
    }
  __syncthreads();
  if (threadIdx.x < kNFeat) {
    atomicMin(&lo[threadIdx.x], slo[threadIdx.x]);
    atomicMax(&hi[threadIdx.x], shi[threadIdx.x]);
  }
}

// ---------------------------------------------------------------- surrogate renderer
// Two passes over the same tiles.  kExact = false renders every tile with the network
// alone and marks (redo[tile] = 1) the tiles where some sample's inputs left the
// training range; kExact = true then renders only the marked tiles again, exact cone
// tracer for those samples, and clears their marks, so every mark is 0 between frames.
// The first pass carries no exact-tracer code (its registers and scalar state serve the
// march, the features and the network alone); the second costs one launch when nothing
// is marked.  Counters: samples and shaded samples from the first pass, exact fallbacks
// from the second (which marches a tile exactly as the first did).
template <typename T, int F, bool kExact>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_nn_render(RayFrame rf, VolGeom g, PyrGeom pg, const void* __restrict__ pyr,
                                                   ConeK K, NnK nk, NnSched sc, const short* __restrict__ wpack,
                                                   const float* __restrict__ bpack, const float4* __restrict__ tf,
                                                   int ntf, float step, float et, float4* __restrict__ out,
                                                   unsigned int* __restrict__ redo, unsigned long long* __restrict__ counters) {
  __shared__ __align__(16) short sW[kPackShorts];
  __shared__ __align__(16) float sB[kNBias];
  __shared__ __align__(16) float sX[4][kWaveScratch];
  extern __shared__ float4 s_tf[];
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  if (kExact && redo[bid] == 0u) return;  // the same word for every thread of the block
  nn_stage(wpack, bpack, tf, ntf, sW, sB, s_tf);
  __syncthreads();
  if (kExact && threadIdx.x == 0) redo[bid] = 0u;  // every thread has read it (barrier above)
  const int tw = 1 << sc.tw_log2;
  const int tiles_x = (rf.W + 2 * tw - 1) / (2 * tw);
  unsigned int cnt[3] = {0, 0, 0};
  const int th = (64 >> sc.tw_log2) * sc.span;
  int tx, ty;
  super_tile(bid, tiles_x, (rf.H + 2 * th - 1) / (2 * th), 2 * tw, 16, tx, ty);
  const bool marked = nn_tile<T, F, kExact>(rf, g, pg, pyr, K, nk, sc, sW, sB, sX[threadIdx.x >> 6], s_tf, ntf, step,
                                            et, out, tx, ty, cnt);
  if (!kExact && marked && (threadIdx.x & 63) == 0) redo[bid] = 1u;
  if (counters) {
    if (!kExact) {
      atomicAdd(&counters[0], (unsigned long long)cnt[0]);
      atomicAdd(&counters[1], (unsigned long long)cnt[1]);
    } else {
      atomicAdd(&counters[2], (unsigned long long)cnt[2]);
    }
  }
}

// ---------------------------------------------------------------- host side
int cone_points(cvr_ctx* c, const cvr_cone_params& P, int n, const float* d_pos, const float* d_dirs,
                const float* d_lights, float* d_occ, float* d_sdw, float* d_feat, hipStream_t st);
int prepare_cone(cvr_ctx* c, const cvr_cone_params& P, hipStream_t st);
ConeK cone_constants(const cvr_ctx* c, const cvr_cone_params& P, const float L[3]);
void light_dir(const RayFrame& rf, const cvr_light* light, float L[3]);
int launch_cone(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st);

static int ensure(void** p, size_t bytes) {
  if (*p) return 0;
  if (hipMalloc(p, bytes) != hipSuccess) {
    *p = nullptr;
    set_error("surrogate: out of device memory");
    return -5;
  }
  return 0;
}

int nn_alloc(cvr_ctx* c) {
  int rc = 0;
  rc |= ensure(reinterpret_cast<void**>(&c->d_nn_w), kNParams * sizeof(float));
  rc |= ensure(reinterpret_cast<void**>(&c->d_nn_g), kNParams * sizeof(float));
  rc |= ensure(reinterpret_cast<void**>(&c->d_nn_m), kNParams * sizeof(float));
  rc |= ensure(reinterpret_cast<void**>(&c->d_nn_v), kNParams * sizeof(float));
  rc |= ensure(reinterpret_cast<void**>(&c->d_nn_pack), kPackShorts * sizeof(short) + kNBias * sizeof(float));
  rc |= ensure(reinterpret_cast<void**>(&c->d_nn_scalar), 64 * sizeof(float));
  return rc ? -5 : 0;
}

int nn_pack(cvr_ctx* c, hipStream_t st) {
  short* wp = static_cast<short*>(c->d_nn_pack);
  float* bp = reinterpret_cast<float*>(wp + kPackShorts);
  hipLaunchKernelGGL(k_mlp_pack, dim3(kNTiles), dim3(64), 0, st, c->d_nn_w, wp, bp);
  CVR_HIP_CHECK(hipGetLastError());
  return 0;
}

// He-uniform initial weights (seeded), zero biases.
static void init_weights(std::vector<float>& w, uint64_t seed) {
  w.assign(kNParams, 0.f);
  uint64_t s = seed ^ 0xA0761D6478BD642Full;
  auto u = [&]() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0);
  };
  const int offs[4] = {kOffW1, kOffW2, kOffW3, kOffW4}, ins[4] = {kNFeat, kH, kH, kH}, outs[4] = {kH, kH, kH, kOut};
  for (int l = 0; l < 4; ++l) {
    const double a = std::sqrt(6.0 / ins[l]) * (l == 3 ? 0.5 : 1.0);
    for (int i = 0; i < outs[l] * ins[l]; ++i) w[offs[l] + i] = float((2.0 * u() - 1.0) * a);
  }
}

// Device buffers of training: the bf16 training packs and the gradient slabs of n samples.
static int train_buffers(cvr_ctx* c, int n) {
  if (!c->d_nn_tpack && hipMalloc(&c->d_nn_tpack, kTrainPackBytes) != hipSuccess) {
    c->d_nn_tpack = nullptr;
    set_error("surrogate: out of device memory (training packs)");
    return -5;
  }
  const size_t need = size_t((n + kTrainChunk - 1) / kTrainChunk) * kSlab;
  if (need > c->nn_slabs_cap) {
    if (c->d_nn_slabs) CVR_HIP_CHECK(hipFree(c->d_nn_slabs));
    c->d_nn_slabs = nullptr;
    c->nn_slabs_cap = 0;
    if (hipMalloc(&c->d_nn_slabs, need * sizeof(float)) != hipSuccess) {
      c->d_nn_slabs = nullptr;
      set_error("surrogate: out of device memory (gradient slabs)");
      return -5;
    }
    c->nn_slabs_cap = need;
  }
  return 0;
}

// The update kernel in one of its modes (0: gradient sum, 1: Adam step, 2: packs only).
static void launch_update(cvr_ctx* c, int nslabs, int mode, float lr, float bc1, float bc2, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_update, dim3((kNParams + 1 + kUpdParams - 1) / kUpdParams), dim3(kUpdParams * kUpdParts), 0, st,
                     c->d_nn_slabs, nslabs, mode,
                     c->d_nn_w, c->d_nn_g, c->d_nn_m, c->d_nn_v, lr, 0.9f, 0.999f, 1e-8f, bc1, bc2,
                     static_cast<unsigned char*>(c->d_nn_tpack), static_cast<short*>(c->d_nn_pack), c->d_nn_scalar);
}

// Gradient of L = sum (y - t)^2 / (2 n) over n samples at the current weights.
int nn_gradient(cvr_ctx* c, const float* d_X, const float* d_T, int n, hipStream_t st) {
  int rc = train_buffers(c, n);
  if (rc) return rc;
  const int chunks = (n + kTrainChunk - 1) / kTrainChunk;
  launch_update(c, 0, 2, 0.f, 1.f, 1.f, st);  // the packs of the current weights
  hipLaunchKernelGGL(k_mlp_train, dim3(chunks), dim3(64 * kTrainWaves), 0, st,
                     static_cast<const unsigned char*>(c->d_nn_tpack), d_X, d_T, n, float(n), c->d_nn_slabs);
  launch_update(c, chunks, 0, 0.f, 1.f, 1.f, st);
  CVR_HIP_CHECK(hipGetLastError());
  return 0;
}

int nn_eval(cvr_ctx* c, const float* d_X, int n, float* d_Y, int packed, hipStream_t st) {
  if (packed) {
    short* wp = static_cast<short*>(c->d_nn_pack);
    float* bp = reinterpret_cast<float*>(wp + kPackShorts);
    hipLaunchKernelGGL(k_mlp_infer, dim3((n + 255) / 256), dim3(256), 0, st, wp, bp, d_X, n, d_Y);
  } else {
    hipLaunchKernelGGL(k_mlp_eval, dim3((n + 63) / 64), dim3(64), 0, st, c->d_nn_w, d_X, n, d_Y);
  }
  CVR_HIP_CHECK(hipGetLastError());
  return 0;
}

int nn_generate(cvr_ctx* c, const cvr_cone_params& P, uint64_t seed, int n, float* d_X, float* d_T, float* d_pos,
                hipStream_t st) {
  int rc = prepare_cone(c, P, st);
  if (rc) return rc;
  float L[3] = {0, 0, 1};
  ConeK K = cone_constants(c, P, L);
  VolGeom g = geom_of(c);
  with_types(c->dtype, c->pyr.fmt, [&](auto t, auto f) {
    using T = typename decltype(t)::type;
    hipLaunchKernelGGL((k_nn_samples<T, decltype(f)::value>), dim3((n + 255) / 256), dim3(256), 0, st, g, c->pyr,
                       c->d_pyr_oct, K, c->d_tf_rgbt, c->tf_size, seed, n, d_X, d_T, d_pos);
  });
  CVR_HIP_CHECK(hipGetLastError());
  return 0;
}

// Generates the training and validation sets with the exact tracer, trains with Adam,
// measures the validation RMSE of the network and of the network-free baseline, and
// records the feature range used to send out-of-range samples to the exact tracer.
int nn_train(cvr_ctx* c, hipStream_t st) {
  const cvr_nn_params& N = c->nn;
  int rc = nn_alloc(c);
  if (rc) return rc;
  const int n = N.n_train, nv = N.n_val;
  if (n < 64 || nv < 64 || N.batch < 64 || N.epochs < 0) {
    set_error("surrogate: n_train, n_val, batch >= 64 and epochs >= 0 required");
    return -1;
  }
  float *d_X = nullptr, *d_T = nullptr, *d_Xv = nullptr, *d_Tv = nullptr, *d_Yv = nullptr;
  unsigned int* d_rng = nullptr;
  auto cleanup = [&]() {
    if (d_X) (void)hipFree(d_X);
    if (d_T) (void)hipFree(d_T);
    if (d_Xv) (void)hipFree(d_Xv);
    if (d_Tv) (void)hipFree(d_Tv);
    if (d_Yv) (void)hipFree(d_Yv);
    if (d_rng) (void)hipFree(d_rng);
  };
  if (hipMalloc(&d_X, size_t(n) * kNFeat * 4) != hipSuccess || hipMalloc(&d_T, size_t(n) * 2 * 4) != hipSuccess ||
      hipMalloc(&d_Xv, size_t(nv) * kNFeat * 4) != hipSuccess || hipMalloc(&d_Tv, size_t(nv) * 2 * 4) != hipSuccess ||
      hipMalloc(&d_Yv, size_t(nv) * 2 * 4) != hipSuccess || hipMalloc(&d_rng, 2 * kNFeat * 4) != hipSuccess) {
    cleanup();
    set_error("surrogate: out of device memory for the training set");
    return -5;
  }
  rc = nn_generate(c, N.cone, N.seed, n, d_X, d_T, nullptr, st);
  if (!rc) rc = nn_generate(c, N.cone, N.seed ^ 0x5851F42D4C957F2Dull, nv, d_Xv, d_Tv, nullptr, st);
  if (rc) { cleanup(); return rc; }
  // feature range
  std::vector<unsigned int> init(2 * kNFeat);
  for (int i = 0; i < kNFeat; ++i) { init[i] = 0x7F800000u; init[kNFeat + i] = 0u; }
  (void)hipMemcpy(d_rng, init.data(), init.size() * 4, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_feat_range, dim3(256), dim3(256), 0, st, d_X, n, d_rng, d_rng + kNFeat);
  std::vector<unsigned int> rng(2 * kNFeat);
  if (hipMemcpyAsync(rng.data(), d_rng, rng.size() * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    cleanup();
    set_error("surrogate: feature range copy failed");
    return -2;
  }
  for (int i = 0; i < kNFeat; ++i) {
    float lo, hi;
    std::memcpy(&lo, &rng[i], 4);
    std::memcpy(&hi, &rng[kNFeat + i], 4);
    c->nn_lo[i] = lo;
    c->nn_hi[i] = hi;
  }
  // weights and optimiser state
  std::vector<float> w0;
  init_weights(w0, N.seed);
  (void)hipMemcpy(c->d_nn_w, w0.data(), kNParams * 4, hipMemcpyHostToDevice);
  (void)hipMemset(c->d_nn_m, 0, kNParams * 4);
  (void)hipMemset(c->d_nn_v, 0, kNParams * 4);
  (void)hipMemset(c->d_nn_g, 0, kNParams * 4);
  const int B = std::min(N.batch, n);
  const int steps_per_epoch = n / B;
  const int total = steps_per_epoch * N.epochs;
  const int chunks = (B + kTrainChunk - 1) / kTrainChunk;  // one workgroup (and gradient slab) each
  rc = train_buffers(c, B);
  if (rc) { cleanup(); return rc; }
  launch_update(c, 0, 2, 0.f, 1.f, 1.f, st);  // packs of the initial weights
  const double b1 = 0.9, b2 = 0.999;
  double bc1 = 1.0, bc2 = 1.0;
  for (int t = 0; t < total; ++t) {
    const int off = (t % steps_per_epoch) * B;
    hipLaunchKernelGGL(k_mlp_train, dim3(chunks), dim3(64 * kTrainWaves), 0, st,
                       static_cast<const unsigned char*>(c->d_nn_tpack), d_X + size_t(off) * kNFeat, d_T + size_t(off) * 2,
                       B, float(B), c->d_nn_slabs);
    bc1 *= b1;
    bc2 *= b2;
    // cosine decay of the learning rate over the run
    const float lr = float(N.lr * 0.5 * (1.0 + std::cos(3.14159265358979 * double(t) / double(std::max(total, 1)))));
    launch_update(c, chunks, 1, lr, float(1.0 - bc1), float(1.0 - bc2), st);
  }
  if (hipGetLastError() != hipSuccess) { cleanup(); set_error("surrogate: training launch failed"); return -2; }
  float last_loss = 0.f;
  (void)hipMemcpyAsync(&last_loss, c->d_nn_scalar, 4, hipMemcpyDeviceToHost, st);
  rc = nn_pack(c, st);
  if (!rc) rc = nn_eval(c, d_Xv, nv, d_Yv, 1, st);
  std::vector<float> Xv(size_t(nv) * kNFeat), Tv(size_t(nv) * 2), Yv(size_t(nv) * 2);
  if (!rc && (hipMemcpyAsync(Xv.data(), d_Xv, Xv.size() * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
              hipMemcpyAsync(Tv.data(), d_Tv, Tv.size() * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
              hipMemcpyAsync(Yv.data(), d_Yv, Yv.size() * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
              hipStreamSynchronize(st) != hipSuccess)) {
    set_error("surrogate: validation copy failed");
    rc = -2;
  }
  cleanup();
  if (rc) return rc;
  // network-free baseline from the same lookups: transmittance products along the
  // axial cone and the ring, weighted like the exact tracer's cones, and along the light
  const cvr_cone_params& P = N.cone;
  const double w_ring = double(P.n_occ_cones - 1) * std::cos(double(P.ring_deg) * 3.14159265358979323846 / 180.0);
  double se[2] = {0, 0}, sb[2] = {0, 0};
  for (int i = 0; i < nv; ++i) {
    const float* x = &Xv[size_t(i) * kNFeat];
    double ta = 1.0, tr = 1.0, base_s = 1.0;
    for (int s = 0; s < kNOcc; ++s) { ta *= 1.0 - x[2 * s]; tr *= 1.0 - x[2 * s + 1]; }
    for (int s = 0; s < kNSdw; ++s) base_s *= 1.0 - x[2 * kNOcc + s];
    const double base_o = (ta + w_ring * tr) / (1.0 + w_ring);
    const double to = double(Tv[i * 2]) + base_o, ts = double(Tv[i * 2 + 1]) + base_s;  // exact values
    const double yo = std::min(std::max(double(Yv[i * 2]) + base_o, 0.0), 1.0);
    const double ys = std::min(std::max(double(Yv[i * 2 + 1]) + base_s, 0.0), 1.0);
    se[0] += (yo - to) * (yo - to);
    se[1] += (ys - ts) * (ys - ts);
    sb[0] += (base_o - to) * (base_o - to);
    sb[1] += (base_s - ts) * (base_s - ts);
  }
  c->nn_report.train_loss = last_loss;
  c->nn_report.val_rmse_occ = std::sqrt(se[0] / nv);
  c->nn_report.val_rmse_sdw = std::sqrt(se[1] / nv);
  c->nn_report.base_rmse_occ = std::sqrt(sb[0] / nv);
  c->nn_report.base_rmse_sdw = std::sqrt(sb[1] / nv);
  c->nn_report.steps = total;
  c->nn_report.exact_fallback = (std::max(c->nn_report.val_rmse_occ, c->nn_report.val_rmse_sdw) > N.rmse_gate) ? 1 : 0;
  c->nn_trained = true;
  c->nn_version++;
  return 0;
}

// Render-time constants of the surrogate: the training range widened by the margin,
// and whether the validation gate (or force_exact) sends every sample to the exact tracer.
NnK nn_consts(cvr_ctx* c) {
  const cvr_nn_params& N = c->nn;
  NnK nk{};
  for (int i = 0; i < kNFeat; ++i) { nk.lo[i] = c->nn_lo[i] - N.ood_margin; nk.hi[i] = c->nn_hi[i] + N.ood_margin; }
  c->nn_report.exact_fallback =
      (std::max(c->nn_report.val_rmse_occ, c->nn_report.val_rmse_sdw) > N.rmse_gate) ? 1 : 0;
  nk.exact_all = (N.force_exact || c->nn_report.exact_fallback) ? 1 : 0;
  return nk;
}

// Work shape of a surrogate frame: the schedule and the number of workgroup tiles.
static NnSched nn_sched(const cvr_nn_params& N) {
  NnSched sc{0, N.shade_batch, N.span};
  while ((1 << sc.tw_log2) < N.tile_w) ++sc.tw_log2;
  return sc;
}
static int nn_tiles(const NnSched& sc, const RayFrame& rf) {
  const int tw = 1 << sc.tw_log2, th = (64 / tw) * sc.span;
  return ((rf.W + 2 * tw - 1) / (2 * tw)) * ((rf.H + 2 * th - 1) / (2 * th));
}

// The frame of a prepared, trained handle: the exact tracer when the validation gate is
// closed, else the two passes of k_nn_render over `redo` (one word per workgroup tile,
// zero on entry and left zero by the second pass).  Allocates nothing.
static int nn_frame(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, unsigned int* redo,
                    unsigned long long* cnt, hipStream_t st) {
  const cvr_nn_params& N = c->nn;
  float L[3];
  light_dir(rf, light, L);
  ConeK K = cone_constants(c, N.cone, L);
  const NnK nk = nn_consts(c);
  if (nk.exact_all) {
    // the validation gate failed: the exact tracer renders the frame
    cvr_cone_params saved = c->cone;
    c->cone = N.cone;
    const int rc = launch_cone(c, rf, light, out, st);
    c->cone = saved;
    return rc;
  }
  VolGeom g = geom_of(c);
  const float step = N.cone.step_size > 0 ? N.cone.step_size : default_step(c);
  const NnSched sc = nn_sched(N);
  const int tiles = nn_tiles(sc, rf);
  const size_t shm = size_t(c->tf_size) * sizeof(float4);
  if (shm > 32 * 1024) { set_error("surrogate renderer: transfer function larger than 2048 entries"); return -4; }
  short* wp = static_cast<short*>(c->d_nn_pack);
  float* bp = reinterpret_cast<float*>(wp + kPackShorts);
  float4* o4 = reinterpret_cast<float4*>(out);
  with_types(c->dtype, c->pyr.fmt, [&](auto t, auto f) {
    using T = typename decltype(t)::type;
    constexpr int FF = decltype(f)::value;
    hipLaunchKernelGGL((k_nn_render<T, FF, false>), dim3(tiles), dim3(256), shm, st, rf, g, c->pyr, c->d_pyr_oct, K,
                       nk, sc, wp, bp, c->d_tf_rgbt, c->tf_size, step, N.cone.et_threshold, o4, redo, cnt);
    hipLaunchKernelGGL((k_nn_render<T, FF, true>), dim3(tiles), dim3(256), shm, st, rf, g, c->pyr, c->d_pyr_oct, K,
                       nk, sc, wp, bp, c->d_tf_rgbt, c->tf_size, step, N.cone.et_threshold, o4, redo, cnt);
  });
  CVR_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_nn(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st) {
  const cvr_nn_params& N = c->nn;
  int rc = prepare_cone(c, N.cone, st);
  if (rc) return rc;
  if (!c->nn_trained) {
    rc = nn_train(c, st);
    if (rc) return rc;
  }
  const int tiles = nn_tiles(nn_sched(N), rf);
  // tile marks for the second pass: zero when allocated, and the second pass clears
  // what the first sets
  if (size_t(tiles) > c->nn_redo_cap) {
    if (c->d_nn_redo) CVR_HIP_CHECK(hipFree(c->d_nn_redo));
    c->d_nn_redo = nullptr;
    c->nn_redo_cap = 0;
    if (hipMalloc(&c->d_nn_redo, size_t(tiles) * sizeof(unsigned int)) != hipSuccess) {
      c->d_nn_redo = nullptr;
      set_error("surrogate: out of device memory");
      return -5;
    }
    CVR_HIP_CHECK(hipMemset(c->d_nn_redo, 0, size_t(tiles) * sizeof(unsigned int)));
    c->nn_redo_cap = size_t(tiles);
  }
  return nn_frame(c, rf, light, out, c->d_nn_redo, c->collect_stats ? c->d_counters : nullptr, st);
}

// ---------------------------------------------------------------------------------
// Allocation-free rendering (cvr_render_noalloc): the tile marks live in a fixed array
// of the library on every GPU instead of a buffer of the handle.  Each GPU holds its
// own copy of `scratch` (256 KiB) until the process ends; host code uses only the
// address hipGetSymbolAddress gives for the current device, never the array's name.
constexpr int kNoAllocMaxTiles = 1 << 16;
static __device__ __align__(256) unsigned int scratch[kNoAllocMaxTiles];
constexpr int kMaxDevices = 64;
static unsigned int* g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

// The device address of `scratch` on `dev`, which must be the current device; looked up
// once per device.
static int scratch_of(int dev, unsigned int** p) {
  if (dev < 0 || dev >= kMaxDevices) { set_error("cvr_render_noalloc: device index above 63"); return CVR_ERR_UNSUPPORTED; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* a = nullptr;
    const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !a) {
      set_error(std::string("cvr_render_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
      return CVR_ERR_HIP;
    }
    g_scratch[dev] = static_cast<unsigned int*>(a);
  }
  *p = g_scratch[dev];
  return CVR_OK;
}

__global__ __launch_bounds__(256) void k_clear_marks(unsigned int* __restrict__ p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0u;
}

int nn_param_count() { return kNParams; }

}  // namespace cvr

using namespace cvr;

extern "C" int cvr_render_noalloc(cvr_handle h, const cvr_camera* cam, const cvr_light* light, int width, int height,
                                  float* out_device, void* hip_stream) {
  if (!h || !cam || !out_device) { set_error("cvr_render_noalloc: null argument"); return CVR_ERR_INVALID; }
  if (h->technique != CVR_CONE_NN) { set_error("cvr_render_noalloc: only CVR_CONE_NN handles"); return CVR_ERR_UNSUPPORTED; }
  if (width < 1 || height < 1 || size_t(width) * height > (size_t(1) << 28)) {
    set_error("cvr_render_noalloc: image size out of range");
    return CVR_ERR_INVALID;
  }
  if (!(cam->fovy_deg > 0 && cam->fovy_deg < 180)) { set_error("cvr_render_noalloc: fovy must be in (0, 180)"); return CVR_ERR_INVALID; }
  if (!h->d_vol || !h->d_tf_rgbt) { set_error("cvr_render_noalloc: volume and transfer function must be set"); return CVR_ERR_STATE; }
  if (!h->nn_trained || !h->d_nn_pack || !cone_ready(h, h->nn.cone)) {
    set_error("cvr_render_noalloc: call cvr_prepare first (pyramid, cone tables and network must be in place)");
    return CVR_ERR_STATE;
  }
  const RayFrame rf = frame_of(cam, width, height);
  const int tiles = nn_tiles(nn_sched(h->nn), rf);
  if (tiles > kNoAllocMaxTiles) {
    set_error("cvr_render_noalloc: the frame needs " + std::to_string(tiles) + " workgroup tiles, above the limit of " +
              std::to_string(kNoAllocMaxTiles) + " (4096 x 4096 pixels at 4 x 16-pixel wave tiles)");
    return CVR_ERR_INVALID;
  }
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) prev = -1;
  if (prev != h->device) CVR_HIP_CHECK(hipSetDevice(h->device));
  struct Restore {
    int d;
    ~Restore() { if (d >= 0) (void)hipSetDevice(d); }
  } restore{prev != h->device ? prev : -1};
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, out_device) != hipSuccess || at.type != hipMemoryTypeDevice) {
    (void)hipGetLastError();
    set_error("cvr_render_noalloc: out_device must be device memory");
    return CVR_ERR_INVALID;
  }
  unsigned int* marks = nullptr;
  int rc = scratch_of(h->device, &marks);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  // the marks must start at zero: a fill kernel (not a memset node) keeps the call capturable
  hipLaunchKernelGGL(k_clear_marks, dim3((tiles + 255) / 256), dim3(256), 0, st, marks, tiles);
  CVR_HIP_CHECK(hipGetLastError());
  return nn_frame(h, rf, light, out_device, marks, nullptr, st);
}
