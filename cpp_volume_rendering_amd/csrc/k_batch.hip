// Batched surrogate rendering (gfx950): many (volume, camera, light) jobs in one
// persistent launch, captured in a HIP graph.
//
// Work items are (job, workgroup tile) pairs.  With 8 queues the tiles of every job are
// cut into 8 bands of tile rows and band x goes to queue x; a workgroup starts on the
// queue of the XCD it runs on (the hardware deals workgroups round-robin over the 8
// XCDs, so blockIdx % 8), which keeps neighbouring tiles of one image in one XCD's L2,
// and steals from the other queues once its own is empty.  Items are ordered volume by
// volume (largest volume first), so a workgroup restages the network and transfer
// function in LDS only when the volume changes.  Each tile is rendered by the same
// device code as the single-frame surrogate renderer.
#include "cvr_nn.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace cvr {

// Per-volume constants (read with scalar loads: the indices are wave-uniform).
struct VolDesc {
  VolGeom g;
  PyrGeom pg;
  const void* pyr;
  ConeK K;  // light direction set per job
  NnK nk;
  const short* wpack;
  const float* bpack;
  const float4* tf;
  int ntf;
  float step, et;
};

struct JobDesc {
  RayFrame rf;
  float L[3];
  int vol;
};

constexpr int kQueues = 8;

// Clears the queue heads and the counters (a kernel, so that it can sit in the graph).
__global__ __launch_bounds__(64) void k_batch_reset(unsigned int* __restrict__ heads,
                                                    unsigned long long* __restrict__ counters) {
  if (threadIdx.x < kQueues) heads[threadIdx.x] = 0u;
  if (threadIdx.x < 4) counters[threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_batch_render(const VolDesc* __restrict__ vols, const JobDesc* __restrict__ jobs,
                                                      const int2* __restrict__ items, const int* __restrict__ qoff,
                                                      int nq, int tiles_x, NnSched sc, size_t img_px,
                                                      float4* __restrict__ out, unsigned int* __restrict__ heads,
                                                      unsigned long long* __restrict__ counters) {
  __shared__ __align__(16) short sW[kPackShorts];
  __shared__ __align__(16) float sB[kNBias];
  __shared__ __align__(16) float sX[4][kWaveScratch];
  __shared__ int s_item;
  extern __shared__ float4 s_tf[];
  const int home = int(blockIdx.x) % nq;
  int loaded = -1;
  unsigned int cnt[3] = {0, 0, 0};
  for (int qi = 0; qi < nq; ++qi) {
    const int q = (home + qi) % nq;
    const int base = qoff[q], n = qoff[q + 1] - base;
    for (;;) {
      if (threadIdx.x == 0) s_item = int(atomicAdd(&heads[q], 1u));
      __syncthreads();
      const int idx = __builtin_amdgcn_readfirstlane(s_item);
      __syncthreads();  // every lane has read s_item before lane 0 may write it again
      if (idx >= n) break;
      const int2 it = items[base + idx];  // (job, tile)
      const int job = __builtin_amdgcn_readfirstlane(it.x), tile = __builtin_amdgcn_readfirstlane(it.y);
      const JobDesc& J = jobs[job];
      const VolDesc& V = vols[J.vol];
      if (J.vol != loaded) {
        // the barrier above guarantees no wave still reads the previous volume's tables
        nn_stage(V.wpack, V.bpack, V.tf, V.ntf, sW, sB, s_tf);
        loaded = J.vol;
        __syncthreads();
      }
      ConeK K = V.K;
      K.L[0] = J.L[0];
      K.L[1] = J.L[1];
      K.L[2] = J.L[2];
      float4* o = out + size_t(job) * img_px;
#define CVR_BATCH_TILE(TT, FF)                                                                                    \
  nn_tile<TT, FF>(J.rf, V.g, V.pg, V.pyr, K, V.nk, sc, sW, sB, sX[threadIdx.x >> 6], s_tf, V.ntf, V.step, V.et, o, \
                  tile % tiles_x, tile / tiles_x, cnt)
      switch (V.g.dtype * 3 + V.pg.fmt) {
        case 0: CVR_BATCH_TILE(uint8_t, 0); break;
        case 1: CVR_BATCH_TILE(uint8_t, 1); break;
        case 2: CVR_BATCH_TILE(uint8_t, 2); break;
        case 3: CVR_BATCH_TILE(uint16_t, 0); break;
        case 4: CVR_BATCH_TILE(uint16_t, 1); break;
        default: CVR_BATCH_TILE(uint16_t, 2); break;
      }
#undef CVR_BATCH_TILE
    }
  }
  if (counters) {
    atomicAdd(&counters[0], (unsigned long long)cnt[0]);
    atomicAdd(&counters[1], (unsigned long long)cnt[1]);
    atomicAdd(&counters[2], (unsigned long long)cnt[2]);
  }
}

// Sum of squared RGB differences over the pixels either image covers (alpha > 0) and
// the count of those values: the image error of a step schedule against the default one.
__global__ __launch_bounds__(256) void k_rgb_sse(const float4* __restrict__ a, const float4* __restrict__ b, size_t n,
                                                 double* __restrict__ acc) {
  double s = 0.0, c = 0.0;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    const float4 x = a[i], y = b[i];
    if (x.w > 0.f || y.w > 0.f) {
      const double dr = double(x.x) - double(y.x), dg = double(x.y) - double(y.y), db = double(x.z) - double(y.z);
      s += dr * dr + dg * dg + db * db;
      c += 3.0;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_down(s, o);
    c += __shfl_down(c, o);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&acc[0], s);
    atomicAdd(&acc[1], c);
  }
}

}  // namespace cvr

using namespace cvr;

// What a batch copied out of one member handle when it built its tables: the device
// buffers its descriptors point to and the versions of the data behind them.  A member
// whose signature changed (a new volume, transfer function, pyramid, network or
// parameters) makes the batch rebuild its tables and graph before the next replay, so a
// replay never reads a freed buffer or a stale network.
struct MemberSig {
  const void* vol_oct;
  const void* pyr_oct;
  const void* tf;
  const void* pack;
  const void* cone_tab;
  uint64_t vol_v, tf_v, nn_v;
  PyrKey pyr_key;
  int pyr_valid;
  cvr_nn_params prm;
};

struct cvr_batch {
  int device = 0, W = 0, H = 0;
  std::vector<cvr_handle> vols;
  std::vector<MemberSig> sigs;  // per member, as of the last table build
  std::vector<cvr_job> jobs;
  cvr_batch_tuning tune{8, 0.5f, kQueues, 2, 1, 0.0, 1.0f, 0.0f, 0.0};
  // device side
  VolDesc* d_vols = nullptr;
  JobDesc* d_jobs = nullptr;
  int2* d_items = nullptr;
  int* d_qoff = nullptr;
  unsigned int* d_heads = nullptr;
  unsigned long long* d_counters = nullptr;
  size_t cap_vols = 0, cap_jobs = 0, cap_items = 0;
  int max_ntf = 0;
  // graph of (reset, render) for the current jobs / tuning / output
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  float* graph_out = nullptr;
  bool dirty = true, tables_dirty = true;
};

namespace {

struct Guard {
  int prev = -1;
  explicit Guard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~Guard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

int grow(void** p, size_t& cap, size_t bytes) {
  if (bytes <= cap && *p) return CVR_OK;
  if (*p) CVR_HIP_CHECK(hipFree(*p));
  *p = nullptr;
  cap = 0;
  if (hipMalloc(p, std::max<size_t>(bytes, 256)) != hipSuccess) {
    *p = nullptr;
    set_error("batch: out of device memory");
    return CVR_ERR_NOMEM;
  }
  cap = bytes;
  return CVR_OK;
}

void drop_graph(cvr_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_out = nullptr;
}

MemberSig sig_of(const cvr_ctx* c) {
  MemberSig m;
  std::memset(&m, 0, sizeof(m));
  m.vol_oct = c->d_vol_oct;
  m.pyr_oct = c->d_pyr_oct;
  m.tf = c->d_tf_rgbt;
  m.pack = c->d_nn_pack;
  m.cone_tab = c->d_cone_tab;
  m.vol_v = c->vol_version;
  m.tf_v = c->tf_version;
  m.pyr_key = c->pyr_key;
  m.pyr_valid = c->pyr_valid ? 1 : 0;
  m.nn_v = c->nn_trained ? c->nn_version : ~0ull;
  std::memcpy(&m.prm, &c->nn, sizeof(cvr_nn_params));
  return m;
}

// True when some member changed since the tables were built.
bool same_sig(const MemberSig& a, const MemberSig& b) {
  return a.vol_oct == b.vol_oct && a.pyr_oct == b.pyr_oct && a.tf == b.tf && a.pack == b.pack &&
         a.cone_tab == b.cone_tab && a.vol_v == b.vol_v && a.tf_v == b.tf_v && a.nn_v == b.nn_v &&
         a.pyr_key == b.pyr_key && a.pyr_valid == b.pyr_valid && std::memcmp(&a.prm, &b.prm, sizeof(a.prm)) == 0;
}
bool members_changed(const cvr_batch* b) {
  if (b->sigs.size() != b->vols.size()) return true;
  for (size_t v = 0; v < b->vols.size(); ++v)
    if (!same_sig(sig_of(b->vols[v]), b->sigs[v])) return true;
  return false;
}

int tw_log2_of(int tile_w) {
  int l = 0;
  while ((1 << l) < tile_w) ++l;
  return l;
}

// Prepares every volume (pyramid, trained network) and uploads the volume and job
// descriptors and the work queues.
int build_tables(cvr_batch* b) {
  const int nv = int(b->vols.size()), nj = int(b->jobs.size());
  std::vector<VolDesc> vd(nv);
  std::vector<size_t> cells(nv);
  b->max_ntf = 0;
  b->sigs.clear();
  for (int v = 0; v < nv; ++v) {
    cvr_ctx* c = b->vols[v];
    int rc = cvr_prepare(c, nullptr);
    if (rc) return rc;
    b->sigs.push_back(sig_of(c));
    const cvr_nn_params& N = c->nn;
    VolDesc& d = vd[v];
    d.g = geom_of(c);
    d.pg = c->pyr;
    d.pyr = c->d_pyr_oct;
    const float L0[3] = {0.f, 0.f, 1.f};
    d.K = cone_constants(c, N.cone, L0);
    d.nk = nn_consts(c);
    d.wpack = static_cast<const short*>(c->d_nn_pack);
    d.bpack = reinterpret_cast<const float*>(static_cast<const short*>(c->d_nn_pack) + kPackShorts);
    d.tf = c->d_tf_rgbt;
    d.ntf = c->tf_size;
    d.step = (N.cone.step_size > 0 ? N.cone.step_size : default_step(c)) * b->tune.step_scale;
    d.et = N.cone.et_threshold;
    b->max_ntf = std::max(b->max_ntf, c->tf_size);
    cells[v] = size_t(c->nx) * c->ny * c->nz;
  }
  if (size_t(b->max_ntf) * sizeof(float4) > 32 * 1024) {
    set_error("batch: transfer functions larger than 2048 entries");
    return CVR_ERR_UNSUPPORTED;
  }
  std::vector<JobDesc> jd(nj);
  for (int j = 0; j < nj; ++j) {
    const cvr_job& J = b->jobs[j];
    jd[j].rf = frame_of(&J.camera, b->W, b->H);
    jd[j].vol = J.volume;
    const double l = std::sqrt(double(J.light.forward[0]) * J.light.forward[0] +
                               double(J.light.forward[1]) * J.light.forward[1] +
                               double(J.light.forward[2]) * J.light.forward[2]);
    for (int i = 0; i < 3; ++i) jd[j].L[i] = float(J.light.forward[i] / l);
  }
  // job order: largest volume first, then by volume, then as given
  std::vector<int> order(nj);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    const int va = b->jobs[a].volume, vc = b->jobs[c].volume;
    if (cells[va] != cells[vc]) return cells[va] > cells[vc];
    return va < vc;
  });
  const int tl = tw_log2_of(b->tune.tile_w), tw = 1 << tl, th = (64 >> tl) * b->tune.span;
  const int tiles_x = (b->W + 2 * tw - 1) / (2 * tw), tiles_y = (b->H + 2 * th - 1) / (2 * th);
  const int nq = b->tune.queues;
  // Every image's tiles in supertile order (super_tile); with 8 queues (one per XCD:
  // workgroup b starts at queue b % 8) queue q takes supertiles q, q + 8, ... of each
  // image, so the 8 XCDs render 8 adjacent supertiles of one image at a time, as the
  // single-frame renderers' block order does.
  std::vector<std::vector<int2>> q(nq);
  for (int j : order)
    for (int b = 0; b < tiles_x * tiles_y; ++b) {
      int tx, ty;
      super_tile(b, tiles_x, tiles_y, 2 * tw, 16, tx, ty);
      q[(b >> 6) % nq].push_back(make_int2(j, ty * tiles_x + tx));
    }
  std::vector<int2> items;
  std::vector<int> qoff(kQueues + 1, 0);
  for (int i = 0; i < nq; ++i) {
    qoff[i] = int(items.size());
    items.insert(items.end(), q[i].begin(), q[i].end());
  }
  for (int i = nq; i <= kQueues; ++i) qoff[i] = int(items.size());
  int rc;
  if ((rc = grow(reinterpret_cast<void**>(&b->d_vols), b->cap_vols, vd.size() * sizeof(VolDesc)))) return rc;
  if ((rc = grow(reinterpret_cast<void**>(&b->d_jobs), b->cap_jobs, jd.size() * sizeof(JobDesc)))) return rc;
  if ((rc = grow(reinterpret_cast<void**>(&b->d_items), b->cap_items, items.size() * sizeof(int2)))) return rc;
  CVR_HIP_CHECK(hipMemcpy(b->d_vols, vd.data(), vd.size() * sizeof(VolDesc), hipMemcpyHostToDevice));
  CVR_HIP_CHECK(hipMemcpy(b->d_jobs, jd.data(), jd.size() * sizeof(JobDesc), hipMemcpyHostToDevice));
  CVR_HIP_CHECK(hipMemcpy(b->d_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
  CVR_HIP_CHECK(hipMemcpy(b->d_qoff, qoff.data(), qoff.size() * sizeof(int), hipMemcpyHostToDevice));
  b->tables_dirty = false;
  return CVR_OK;
}

int build_graph(cvr_batch* b, float* out) {
  drop_graph(b);
  int ncu = 0;
  CVR_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device));
  const int grid = std::max(1, ncu * b->tune.workgroups_per_cu);
  const int tl = tw_log2_of(b->tune.tile_w), tw = 1 << tl;
  const int tiles_x = (b->W + 2 * tw - 1) / (2 * tw);
  NnSched sc{tl, b->tune.shade_batch, b->tune.span};
  const size_t shm = size_t(b->max_ntf) * sizeof(float4);
  CVR_HIP_CHECK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, b->cap, b->d_heads, b->d_counters);
  hipLaunchKernelGGL(k_batch_render, dim3(grid), dim3(256), shm, b->cap, b->d_vols, b->d_jobs, b->d_items, b->d_qoff,
                     b->tune.queues, tiles_x, sc, size_t(b->W) * b->H, reinterpret_cast<float4*>(out), b->d_heads,
                     b->d_counters);
  const hipError_t le = hipGetLastError();
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(b->cap, &g);
  if (le != hipSuccess || ce != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    set_error(std::string("batch: graph capture failed: ") + hipGetErrorString(le != hipSuccess ? le : ce));
    return CVR_ERR_HIP;
  }
  b->graph = g;
  CVR_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->graph_out = out;
  b->dirty = false;
  return CVR_OK;
}

int check_tuning(const cvr_batch_tuning& t) {
  if ((t.tile_w != 4 && t.tile_w != 8 && t.tile_w != 16 && t.tile_w != 32) || !(t.shade_batch >= 0.f && t.shade_batch <= 1.f) ||
      (t.queues != 1 && t.queues != kQueues) || t.workgroups_per_cu < 1 || t.workgroups_per_cu > 4 ||
      (t.span != 1 && t.span != 2 && t.span != 4 && t.span != 8) || !(t.step_scale >= 0.25f && t.step_scale <= 4.f) ||
      !(t.max_step_rmse >= 0.f && t.max_step_rmse <= 1.f)) {
    set_error("batch tuning: tile_w in {4,8,16,32}, shade_batch in [0,1], queues 1 or 8, workgroups_per_cu in [1,4], "
              "span 1, 2, 4 or 8, step_scale in [0.25, 4], max_step_rmse in [0, 1]");
    return CVR_ERR_INVALID;
  }
  return CVR_OK;
}

// Step search of cvr_batch_autotune on the jobs currently set: renders them at
// step_scale 1 as the reference, then at each larger step_scale in turn up to the first
// whose images differ from the reference by more than max_step_rmse (the error grows
// with the step), and keeps the fastest admitted step.
template <class TimeOf>
int tune_step(cvr_batch* b, float* out, hipStream_t st, TimeOf& time_of, cvr_batch_tuning& cur, double& cur_ms) {
  const size_t px = b->jobs.size() * size_t(b->W) * b->H;
  float4* ref = nullptr;
  double* acc = nullptr;
  if (hipMalloc(&ref, px * sizeof(float4)) != hipSuccess || hipMalloc(&acc, 2 * sizeof(double)) != hipSuccess) {
    if (ref) (void)hipFree(ref);
    set_error("cvr_batch_autotune: out of device memory for the step search");
    return CVR_ERR_NOMEM;
  }
  struct Free {
    void* p[2];
    ~Free() {
      for (void* q : p) (void)hipFree(q);
    }
  } fr{{ref, acc}};
  const float scales[] = {1.f, 1.25f, 1.5f, 2.f, 2.5f, 3.f};
  cvr_batch_tuning t = cur;
  t.step_scale = scales[0];
  t.step_rmse = 0.0;
  int rc = time_of(t, &cur_ms);
  if (rc) return rc;
  cur = t;
  CVR_HIP_CHECK(hipMemcpyAsync(ref, out, px * sizeof(float4), hipMemcpyDeviceToDevice, st));
  const int grid = int(std::min<size_t>((px + 255) / 256, 4096));
  for (int i = 1; i < int(sizeof(scales) / sizeof(scales[0])); ++i) {
    t.step_scale = scales[i];
    double ms = 0;
    if ((rc = time_of(t, &ms))) return rc;
    CVR_HIP_CHECK(hipMemsetAsync(acc, 0, 2 * sizeof(double), st));
    hipLaunchKernelGGL(k_rgb_sse, dim3(grid), dim3(256), 0, st, reinterpret_cast<const float4*>(out), ref, px, acc);
    CVR_HIP_CHECK(hipGetLastError());
    double h[2] = {0.0, 0.0};
    CVR_HIP_CHECK(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, st));
    CVR_HIP_CHECK(hipStreamSynchronize(st));
    const double rmse = h[1] > 0 ? std::sqrt(h[0] / h[1]) : 0.0;
    if (rmse > double(cur.max_step_rmse)) break;
    if (ms < cur_ms) {
      cur = t;
      cur.step_rmse = rmse;
      cur_ms = ms;
    }
  }
  return CVR_OK;
}

}  // namespace

extern "C" {

int cvr_batch_create(int device, int width, int height, cvr_batch_handle* out) {
  if (!out) { set_error("cvr_batch_create: out is null"); return CVR_ERR_INVALID; }
  *out = nullptr;
  if (width < 1 || height < 1 || size_t(width) * height > (size_t(1) << 26)) {
    set_error("cvr_batch_create: image size out of range");
    return CVR_ERR_INVALID;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    set_error("cvr_batch_create: no such HIP device");
    return CVR_ERR_INVALID;
  }
  Guard gd(device);
  cvr_batch* b = new cvr_batch();
  b->device = device;
  b->W = width;
  b->H = height;
  if (hipMalloc(&b->d_heads, kQueues * sizeof(unsigned int)) != hipSuccess ||
      hipMalloc(&b->d_counters, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&b->d_qoff, (kQueues + 1) * sizeof(int)) != hipSuccess ||
      hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess) {
    cvr_batch_destroy(b);
    set_error("cvr_batch_create: HIP allocation failed");
    return CVR_ERR_HIP;
  }
  *out = b;
  return CVR_OK;
}

void cvr_batch_destroy(cvr_batch_handle b) {
  if (!b) return;
  Guard gd(b->device);
  (void)hipDeviceSynchronize();
  drop_graph(b);
  for (void* p : {static_cast<void*>(b->d_vols), static_cast<void*>(b->d_jobs), static_cast<void*>(b->d_items),
                  static_cast<void*>(b->d_qoff), static_cast<void*>(b->d_heads), static_cast<void*>(b->d_counters)})
    if (p) (void)hipFree(p);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  delete b;
}

int cvr_batch_add_volume(cvr_batch_handle b, cvr_handle h, int* index) {
  if (!b || !h) { set_error("cvr_batch_add_volume: null argument"); return CVR_ERR_INVALID; }
  if (h->technique != CVR_CONE_NN) { set_error("cvr_batch_add_volume: the handle must be CVR_CONE_NN"); return CVR_ERR_UNSUPPORTED; }
  if (h->device != b->device) { set_error("cvr_batch_add_volume: handle on another device"); return CVR_ERR_INVALID; }
  if (!h->d_vol || !h->d_tf_rgbt) { set_error("cvr_batch_add_volume: volume and transfer function must be set"); return CVR_ERR_STATE; }
  if (b->vols.size() >= 64) { set_error("cvr_batch_add_volume: at most 64 volumes"); return CVR_ERR_INVALID; }
  b->vols.push_back(h);
  if (index) *index = int(b->vols.size()) - 1;
  b->tables_dirty = b->dirty = true;
  return CVR_OK;
}

int cvr_batch_set_jobs(cvr_batch_handle b, int n, const cvr_job* jobs) {
  if (!b || (n > 0 && !jobs) || n < 0) { set_error("cvr_batch_set_jobs: bad arguments"); return CVR_ERR_INVALID; }
  if (size_t(n) * (size_t(b->W) * b->H) > (size_t(1) << 32)) { set_error("cvr_batch_set_jobs: batch too large"); return CVR_ERR_INVALID; }
  for (int i = 0; i < n; ++i) {
    if (jobs[i].volume < 0 || jobs[i].volume >= int(b->vols.size())) {
      set_error("cvr_batch_set_jobs: volume index out of range");
      return CVR_ERR_INVALID;
    }
    const float* f = jobs[i].light.forward;
    if (!(f[0] * f[0] + f[1] * f[1] + f[2] * f[2] > 0.f)) {
      set_error("cvr_batch_set_jobs: light forward must be nonzero");
      return CVR_ERR_INVALID;
    }
    if (!(jobs[i].camera.fovy_deg > 0 && jobs[i].camera.fovy_deg < 180)) {
      set_error("cvr_batch_set_jobs: fovy must be in (0, 180)");
      return CVR_ERR_INVALID;
    }
  }
  b->jobs.assign(jobs, jobs + n);
  b->tables_dirty = b->dirty = true;
  return CVR_OK;
}

int cvr_batch_set_tuning(cvr_batch_handle b, const cvr_batch_tuning* t) {
  if (!b || !t) { set_error("cvr_batch_set_tuning: null argument"); return CVR_ERR_INVALID; }
  int rc = check_tuning(*t);
  if (rc) return rc;
  const bool items_change = t->tile_w != b->tune.tile_w || t->queues != b->tune.queues || t->span != b->tune.span ||
                            t->step_scale != b->tune.step_scale;
  b->tune = *t;
  if (items_change) b->tables_dirty = true;
  b->dirty = true;
  return CVR_OK;
}

int cvr_batch_get_tuning(cvr_batch_handle b, cvr_batch_tuning* t) {
  if (!b || !t) { set_error("cvr_batch_get_tuning: null argument"); return CVR_ERR_INVALID; }
  *t = b->tune;
  return CVR_OK;
}

int cvr_batch_render(cvr_batch_handle b, float* out, void* hip_stream) {
  if (!b || !out) { set_error("cvr_batch_render: null argument"); return CVR_ERR_INVALID; }
  if (b->jobs.empty()) { set_error("cvr_batch_render: no jobs"); return CVR_ERR_STATE; }
  Guard gd(b->device);
  int rc;
  if (!b->tables_dirty && members_changed(b)) {
    // a member's buffers or network changed: wait for replays that may still read the
    // old ones before the tables are rebuilt
    CVR_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(hip_stream)));
    b->tables_dirty = true;
  }
  if (b->tables_dirty) {
    if ((rc = build_tables(b))) return rc;
    b->dirty = true;
  }
  if (b->dirty || out != b->graph_out)
    if ((rc = build_graph(b, out))) return rc;
  CVR_HIP_CHECK(hipGraphLaunch(b->exec, static_cast<hipStream_t>(hip_stream)));
  return CVR_OK;
}

int cvr_batch_autotune(cvr_batch_handle b, float* out, int stride, void* hip_stream, cvr_batch_tuning* best) {
  if (!b || !out || stride < 1) { set_error("cvr_batch_autotune: bad arguments"); return CVR_ERR_INVALID; }
  if (b->jobs.empty()) { set_error("cvr_batch_autotune: no jobs"); return CVR_ERR_STATE; }
  Guard gd(b->device);
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  const cvr_batch_tuning start = b->tune;
  const std::vector<cvr_job> all = b->jobs;
  std::vector<cvr_job> sub;
  for (size_t i = 0; i < all.size(); i += size_t(stride)) sub.push_back(all[i]);
  b->jobs = sub;
  b->tables_dirty = b->dirty = true;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CVR_HIP_CHECK(hipEventCreate(&e0));
  CVR_HIP_CHECK(hipEventCreate(&e1));
  auto time_of = [&](const cvr_batch_tuning& t, double* ms) -> int {
    int rc = cvr_batch_set_tuning(b, &t);
    if (rc) return rc;
    if ((rc = cvr_batch_render(b, out, st))) return rc;  // warm-up (and graph build)
    CVR_HIP_CHECK(hipEventRecord(e0, st));
    if ((rc = cvr_batch_render(b, out, st))) return rc;
    CVR_HIP_CHECK(hipEventRecord(e1, st));
    CVR_HIP_CHECK(hipEventSynchronize(e1));
    float f = 0.f;
    CVR_HIP_CHECK(hipEventElapsedTime(&f, e0, e1));
    *ms = f;
    return CVR_OK;
  };
  // the step schedule first (its images against the step_scale 1 images of the same
  // jobs), then coordinate descent over tile width, shading batch, queues, workgroups
  // per CU and span at the chosen step
  cvr_batch_tuning cur = b->tune;
  cur.step_rmse = 0.0;
  double cur_ms = 0;
  int rc = CVR_OK;
  if (cur.max_step_rmse > 0.f) rc = tune_step(b, out, st, time_of, cur, cur_ms);
  else rc = time_of(cur, &cur_ms);
  const int tws[] = {4, 8, 16, 32};
  const float sbs[] = {0.f, 0.25f, 0.5f, 0.75f, 1.f};
  const int qs[] = {1, kQueues};
  const int wgs[] = {1, 2, 3};
  const int spans[] = {1, 2, 4, 8};
  for (int pass = 0; pass < 5 && rc == CVR_OK; ++pass) {
    const int nopt = pass == 0 ? 4 : pass == 1 ? 5 : pass == 2 ? 2 : pass == 3 ? 3 : 4;
    for (int o = 0; o < nopt && rc == CVR_OK; ++o) {
      cvr_batch_tuning t = cur;
      if (pass == 0) t.tile_w = tws[o];
      if (pass == 1) t.shade_batch = sbs[o];
      if (pass == 2) t.queues = qs[o];
      if (pass == 3) t.workgroups_per_cu = wgs[o];
      if (pass == 4) t.span = spans[o];
      if (t.tile_w == cur.tile_w && t.shade_batch == cur.shade_batch && t.queues == cur.queues &&
          t.workgroups_per_cu == cur.workgroups_per_cu && t.span == cur.span)
        continue;
      double ms = 0;
      rc = time_of(t, &ms);
      if (rc == CVR_OK && ms < cur_ms) {
        cur = t;
        cur_ms = ms;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->jobs = all;
  b->tables_dirty = b->dirty = true;
  if (rc) {
    (void)cvr_batch_set_tuning(b, &start);
    return rc;
  }
  cur.ms = cur_ms;
  b->tune = cur;
  if (best) *best = cur;
  return CVR_OK;
}

int cvr_batch_get_stats(cvr_batch_handle b, cvr_stats* out) {
  if (!b || !out) { set_error("cvr_batch_get_stats: null argument"); return CVR_ERR_INVALID; }
  Guard gd(b->device);
  CVR_HIP_CHECK(hipDeviceSynchronize());
  unsigned long long c[4];
  CVR_HIP_CHECK(hipMemcpy(c, b->d_counters, sizeof(c), hipMemcpyDeviceToHost));
  out->samples = c[0];
  out->shaded_samples = c[1];
  out->fallback_samples = c[2];
  out->last_ms = 0.0;
  return CVR_OK;
}

}  // extern "C"
