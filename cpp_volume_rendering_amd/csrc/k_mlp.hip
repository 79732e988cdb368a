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
// Training (k_mlp_train): one wavefront per workgroup walks `chunks` groups of 64
// samples, keeps the activations of one group in LDS, runs the backward GEMMs
// (W^T dY and dY H^T, operands gathered from LDS with transposed strides) and
// accumulates the weight gradients in LDS; one atomic flush per workgroup.
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

// Gradients of one layer: dW += G_out H_in^T, db += row sums of G_out, and (if gIn)
// gIn = (W^T G_out) masked by H_in > 0.
__device__ void layer_bwd(const float* __restrict__ W, int in, int out, const float* gOut, const float* hIn,
                          float* gIn, float* sGradW, float* sGradB) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  if (lane < out) {
    float s = 0.f;
    for (int j = 0; j < 64; ++j) s += gOut[lane * 64 + j];
    sGradB[lane] += s;
  }
  for (int mt = 0; mt < out / 16; ++mt)
    for (int nt = 0; nt < in / 16; ++nt) {
      f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
      acc = tile_gemm(gOut, 64, 1, hIn, 1, 64, mt * 16, nt * 16, 64, acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) sGradW[(mt * 16 + 4 * q + e) * in + nt * 16 + r] += acc[e];
    }
  if (!gIn) return;
  for (int mt = 0; mt < in / 16; ++mt)
    for (int nt = 0; nt < 4; ++nt) {
      f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
      acc = tile_gemm(W, 1, in, gOut, 64, 1, mt * 16, nt * 16, out, acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = mt * 16 + 4 * q + e, s = nt * 16 + r;
        gIn[i * 64 + s] = hIn[i * 64 + s] > 0.f ? acc[e] : 0.f;
      }
    }
}

// X: [n][16] features, T: [n][2] targets.  grad (global, zeroed by the caller) +=
// d/dparams of sum (y - t)^2 / (2 * norm_n); loss (global) += that sum.
__global__ __launch_bounds__(64) void k_mlp_train(const float* __restrict__ prm, const float* __restrict__ X,
                                                  const float* __restrict__ Tg, int n, int chunks, float norm_n,
                                                  float* __restrict__ grad, float* __restrict__ loss) {
  extern __shared__ float lds[];
  float* sX = lds;               // 16 x 64
  float* sH1 = sX + 16 * 64;     // 64 x 64
  float* sH2 = sH1 + 64 * 64;
  float* sH3 = sH2 + 64 * 64;
  float* sY = sH3 + 64 * 64;     // 16 x 64
  float* sG1 = sY + 16 * 64;     // 64 x 64
  float* sG2 = sG1 + 64 * 64;    // 64 x 64
  float* sGrad = sG2 + 64 * 64;  // kNParams
  const int lane = threadIdx.x;
  for (int i = lane; i < kNParams; i += 64) sGrad[i] = 0.f;
  float lsum = 0.f;
  const float inv_n = 1.0f / norm_n;
  for (int c = 0; c < chunks; ++c) {
    const int s0 = (blockIdx.x * chunks + c) * 64;
    if (s0 >= n) break;
    __syncthreads();
    const int s = s0 + lane;
    const bool valid = s < n;
    for (int f = 0; f < kNFeat; ++f) sX[f * 64 + lane] = valid ? X[size_t(s) * kNFeat + f] : 0.f;
    __syncthreads();
    layer_fwd(prm + kOffW1, prm + kOffB1, kNFeat, kH, sX, sH1, true);
    __syncthreads();
    layer_fwd(prm + kOffW2, prm + kOffB2, kH, kH, sH1, sH2, true);
    __syncthreads();
    layer_fwd(prm + kOffW3, prm + kOffB3, kH, kH, sH2, sH3, true);
    __syncthreads();
    layer_fwd(prm + kOffW4, prm + kOffB4, kH, kOut, sH3, sY, false);
    __syncthreads();
    // dL/dy for the 2 used outputs of valid samples
    for (int o = 0; o < kOut; ++o) {
      float g = 0.f;
      if (o < 2 && valid) {
        const float e = sY[o * 64 + lane] - Tg[size_t(s) * 2 + o];
        lsum += e * e;
        g = e * inv_n;
      }
      sG1[o * 64 + lane] = g;
    }
    __syncthreads();
    layer_bwd(prm + kOffW4, kH, kOut, sG1, sH3, sG2, sGrad + kOffW4, sGrad + kOffB4);
    __syncthreads();
    layer_bwd(prm + kOffW3, kH, kH, sG2, sH2, sG1, sGrad + kOffW3, sGrad + kOffB3);
    __syncthreads();
    layer_bwd(prm + kOffW2, kH, kH, sG1, sH1, sG2, sGrad + kOffW2, sGrad + kOffB2);
    __syncthreads();
    layer_bwd(prm + kOffW1, kNFeat, kH, sG2, sX, nullptr, sGrad + kOffW1, sGrad + kOffB1);
  }
  __syncthreads();
  for (int i = lane; i < kNParams; i += 64) atomicAdd(&grad[i], sGrad[i]);
  // wave sum of the squared errors
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off, 64);
  if (lane == 0) atomicAdd(loss, lsum * 0.5f * inv_n);
}

// Adam (Kingma & Ba 2015) on the fp32 master parameters; clears the gradient.
__global__ void k_adam(float* __restrict__ prm, float* __restrict__ grad, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = grad[i];
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  prm[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  grad[i] = 0.f;
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

constexpr size_t kTrainLds = (16 * 64 * 2 + 64 * 64 * 5 + kNParams) * sizeof(float);

// The training kernel keeps one 64-sample group's activations and the whole gradient
// in LDS (~129 KiB of the 160 KiB gfx950 gives a workgroup).
static int train_kernel_attr() {
  static int done = 0;
  if (!done) {
    CVR_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_mlp_train),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(kTrainLds)));
    done = 1;
  }
  return 0;
}

// Gradient of L = sum (y - t)^2 / (2 n) over n samples at the current weights.
int nn_gradient(cvr_ctx* c, const float* d_X, const float* d_T, int n, hipStream_t st) {
  CVR_HIP_CHECK(hipMemsetAsync(c->d_nn_g, 0, kNParams * sizeof(float), st));
  CVR_HIP_CHECK(hipMemsetAsync(c->d_nn_scalar, 0, sizeof(float), st));
  const int chunks = 4;
  const int blocks = (n + 64 * chunks - 1) / (64 * chunks);
  const size_t shm = kTrainLds;
  int rc = train_kernel_attr();
  if (rc) return rc;
  hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(64), shm, st, c->d_nn_w, d_X, d_T, n, chunks, float(n),
                     c->d_nn_g, c->d_nn_scalar);
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
  // 64-sample groups per workgroup: enough workgroups to fill the 256 CUs
  const int chunks = std::max(1, std::min(4, B / (64 * 256)));
  const size_t shm = kTrainLds;
  rc = train_kernel_attr();
  if (rc) { cleanup(); return rc; }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double bc1 = 1.0, bc2 = 1.0;
  for (int t = 0; t < total; ++t) {
    const int off = (t % steps_per_epoch) * B;
    (void)hipMemsetAsync(c->d_nn_scalar, 0, sizeof(float), st);
    hipLaunchKernelGGL(k_mlp_train, dim3((B + 64 * chunks - 1) / (64 * chunks)), dim3(64), shm, st, c->d_nn_w,
                       d_X + size_t(off) * kNFeat, d_T + size_t(off) * 2, B, chunks, float(B), c->d_nn_g,
                       c->d_nn_scalar);
    bc1 *= b1;
    bc2 *= b2;
    // cosine decay of the learning rate over the run
    const float lr = float(N.lr * 0.5 * (1.0 + std::cos(3.14159265358979 * double(t) / double(std::max(total, 1)))));
    hipLaunchKernelGGL(k_adam, dim3((kNParams + 255) / 256), dim3(256), 0, st, c->d_nn_w, c->d_nn_g, c->d_nn_m,
                       c->d_nn_v, kNParams, lr, b1, b2, eps, float(1.0 - bc1), float(1.0 - bc2));
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
