// Gather-peak microbenchmark: the ceiling of the march kernels' trilinear reads.
//
// Every read of the renderers is one 32-byte record (a corner block: the 8 texels of a
// GL_LINEAR lookup) at an address that depends on the ray, so their rate is bounded by
// how fast the memory hierarchy returns independent 32-byte records to the lanes that
// ask for them, from the level that holds the working set.  k_probe_gather measures
// exactly that: every lane reads records at pseudo-random indices (a per-lane
// xorshift sequence, so no read waits on another) from a working set of a chosen
// size, eight reads in flight per lane, each as the two 16-byte loads a corner-block
// read issues.  Working sets of 16 KiB stay in each CU's L1, 2 MiB in each XCD's L2,
// 64 MiB in the shared MALL (256 MB) and 4 GiB in HBM only.  The rate counts 32 bytes
// per read, the same algorithmic bytes the renderers' roofline figures count.
#include "cvr_ctx.hpp"
#include <algorithm>
#include <string>

namespace cvr {

__global__ __launch_bounds__(256) void k_probe_gather(const uint4* __restrict__ recs, uint32_t nmask, int iters,
                                                      uint32_t seed, uint4* __restrict__ sink) {
  uint32_t h = ((blockIdx.x * 256u + threadIdx.x) * 0x9E3779B1u ^ seed) | 1u;  // nonzero
  uint4 acc = make_uint4(0u, 0u, 0u, 0u);
  for (int it = 0; it < iters; ++it) {
    uint4 a[8], b[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      h ^= h << 13;  // xorshift32: every bit of the index is well mixed
      h ^= h >> 17;
      h ^= h << 5;
      const uint32_t i = h & nmask;
      a[j] = recs[2 * size_t(i)];
      b[j] = recs[2 * size_t(i) + 1];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      acc.x ^= a[j].x ^ b[j].x;
      acc.y ^= a[j].y ^ b[j].y;
      acc.z ^= a[j].z ^ b[j].z;
      acc.w ^= a[j].w ^ b[j].w;
    }
  }
  // the buffer holds zeros: never true, but the reads cannot be dropped
  if ((acc.x | acc.y | acc.z | acc.w) == 0xFFFFFFFFu) sink[0] = acc;
}

}  // namespace cvr

using namespace cvr;

// working_set_bytes: rounded down to a power of two of 32-byte records (>= 1024 bytes,
// <= 16 GiB).  Returns the best of `reps` timed launches (after one warm-up) as GB/s of
// 32-byte reads, and the reads per launch.
extern "C" int cvr_probe_gather(int device, size_t working_set_bytes, int reps, double* gbps, double* reads_per_launch,
                                void* hip_stream) {
  if (!gbps || reps < 1 || working_set_bytes < 1024 || working_set_bytes > (size_t(1) << 34)) {
    set_error("cvr_probe_gather: working set in [1 KiB, 16 GiB], reps >= 1 and an output required");
    return CVR_ERR_INVALID;
  }
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) prev = -1;
  if (prev != device) CVR_HIP_CHECK(hipSetDevice(device));
  struct Restore {
    int d;
    ~Restore() { if (d >= 0) (void)hipSetDevice(d); }
  } restore{prev != device ? prev : -1};
  size_t n = 1;
  while (n * 2 * 32 <= working_set_bytes) n *= 2;
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  uint4* d = nullptr;
  if (hipMalloc(&d, n * 32 + 16) != hipSuccess) {
    set_error("cvr_probe_gather: out of device memory");
    return CVR_ERR_NOMEM;
  }
  struct Free {
    void* p;
    ~Free() { (void)hipFree(p); }
  } fr{d};
  CVR_HIP_CHECK(hipMemsetAsync(d, 0, n * 32 + 16, st));
  int ncu = 0;
  CVR_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
  const int blocks = std::max(1, ncu) * 8;  // 8 workgroups of 4 waves per CU
  const int iters = 64;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CVR_HIP_CHECK(hipEventCreate(&e0));
  CVR_HIP_CHECK(hipEventCreate(&e1));
  double best = 0;
  int rc = CVR_OK;
  for (int r = 0; r <= reps && rc == CVR_OK; ++r) {
    if (hipEventRecord(e0, st) != hipSuccess) rc = CVR_ERR_HIP;
    hipLaunchKernelGGL(k_probe_gather, dim3(blocks), dim3(256), 0, st, d, uint32_t(n - 1), iters, 0x1234567u + r,
                       d + 2 * n);
    if (hipGetLastError() != hipSuccess || hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
      rc = CVR_ERR_HIP;
    float ms = 0.f;
    if (rc == CVR_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = CVR_ERR_HIP;
    const double reads = double(blocks) * 256.0 * iters * 8.0;
    if (rc == CVR_OK && r > 0 && ms > 0) best = std::max(best, reads * 32.0 / (double(ms) * 1e-3) / 1e9);
    if (reads_per_launch) *reads_per_launch = reads;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    set_error("cvr_probe_gather: HIP error while timing");
    return rc;
  }
  *gbps = best;
  return CVR_OK;
}
