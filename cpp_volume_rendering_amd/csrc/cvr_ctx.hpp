// Technique instance state behind a cvr_handle.
#pragma once
#include "cvr_common.hpp"
#include "cvr_phong.hpp"
#include "../../include/cvr.h"
#include <vector>

namespace cvr {
constexpr int kMaxLevels = 12;
// Mip chain of the extinction pyramid: level dims and offsets into one fp32 buffer.
struct PyrGeom {
  int levels;
  int fmt;  // storage of the corner blocks: 0 fp32 (2 x float4), 1 fp16, 2 bf16 (8 x 16 bit)
  int dims[kMaxLevels][3];
  size_t off[kMaxLevels];   // level offsets in the linear fp32 copy
  size_t boff[kMaxLevels];  // level offsets (in records) in the brick-ordered corner blocks
  int nb[kMaxLevels][2];    // bricks along x and y per level
  float ts[kMaxLevels][3];  // texel coordinate per world unit (dims / G) per level
  float dmax[kMaxLevels][3];  // dims - 1 per level: the last texel index as a float
  float G[3];
  float one;  // 1.0f, read as a kernel argument: a multiplier the compiler cannot fold
  int wide;   // some level's corner blocks span 4 GiB or more: 64-bit record offsets
};
// One box of the EBS directional-light cone chain (offsets from the sample, the
// CUT_WHEN_AWAY_FROM_VOLUME coordinate apos + interval, 1 / box volume in voxels).
struct ShadowRow {
  float o1[3], o2[3];
  float next;
  float inv_vq;
};
static_assert(sizeof(ShadowRow) == 32, "two float4 per row");
// Inputs a built pyramid depends on, compared field by field before reuse.
struct PyrKey {
  uint64_t vol_v, tf_v;
  int base[3];
  uint32_t sigma_bits;  // bit pattern of sigma0
  int nlev, fmt;
  bool operator==(const PyrKey& o) const {
    return vol_v == o.vol_v && tf_v == o.tf_v && base[0] == o.base[0] && base[1] == o.base[1] &&
           base[2] == o.base[2] && sigma_bits == o.sigma_bits && nlev == o.nlev && fmt == o.fmt;
  }
};
// Section table of one directional-occlusion cone: rows (interval, level, d_integral,
// amplitude), section counts with 1/3/7 Gaussians, sub-cone axes.
struct DosTable {
  std::vector<float4> rows;
  int n[3] = {0, 0, 0};
  float axes[10][3] = {};
  float adj7 = 0.f, init = 0.f, w = 0.f;
};
}  // namespace cvr

struct cvr_ctx {
  int technique = 0;
  int device = 0;

  // volume
  int nx = 0, ny = 0, nz = 0, dtype = 0;
  float sx = 1, sy = 1, sz = 1;
  void* d_vol = nullptr;
  void* d_vol_oct = nullptr;  // corner blocks of the voxels, 8 x u8 / 8 x u16 (k_quad.hip)

  // transfer function (host FP64 table as TransferFunction1D::BuildLinear builds it)
  int tf_max_density = 255;
  bool tf_ext_input = false;
  std::vector<double> tf_tab;  // (max_density+1) * 4
  float4* d_tf_rgbt = nullptr; // GenerateTexture_1D_RGBt values
  float4* d_tf_rgba = nullptr; // GenerateTexture_1D_RGBA values
  int tf_size = 0;
  uint64_t tf_version = 0, vol_version = 0;

  // parameters (one of these is live, chosen by `technique`)
  cvr_rc1p_params rc1p{};
  cvr_ebs_params ebs{};
  cvr_gt_params gt{};
  cvr_dos_params dos{};
  cvr_cone_params cone{};
  cvr_nn_params nn{};

  // EBS summed-area table, two-float (hi, lo) per cell
  float2* d_sat = nullptr;
  void* d_sat_oct = nullptr;  // corner blocks of the SAT, 64 B per cell (k_sat.hip)
  float* d_ext_lut = nullptr;  // per raw value extinction (GetExtN), 256 or 65536 entries
  uint64_t sat_built_for = ~0ull;
  uint64_t sat_oct_built_for = ~0ull;  // d_sat_oct holds the blocks of this SAT

  // Gaussian extinction pyramid (cone tracers, surrogate)
  float* d_pyr = nullptr;
  void* d_pyr_oct = nullptr;   // corner blocks of every level (PyrGeom::fmt, k_quad.hip)
  cvr::PyrGeom pyr{};
  cvr::PyrKey pyr_key{};
  bool pyr_valid = false;  // d_pyr / d_pyr_oct hold the pyramid of pyr_key
  void* d_cone_tab = nullptr;  // surrogate input taps + cone step schedules
  bool cone_tab_valid = false; // d_cone_tab holds the table of cone_tab_for over pyr_key
  cvr_cone_params cone_tab_for{};
  cvr::PyrKey cone_tab_pyr{};  // the pyramid the table's level geometry describes

  // directional occlusion shading: section tables (occlusion rows, then shadow rows)
  cvr::DosTable dos_occ, dos_sdw;
  float4* d_dos_rows = nullptr;
  size_t dos_rows_cap = 0;

  // cone ground truth: sampled ray directions (occlusion, then shadow)
  std::vector<float> gt_occ, gt_sdw;
  float* d_gt_rays = nullptr;
  cvr_gt_params gt_rays_for{};
  int cone_n_steps = 0;

  // surrogate network: fp32 master weights, gradient, Adam moments, packed bf16 copy
  float *d_nn_w = nullptr, *d_nn_g = nullptr, *d_nn_m = nullptr, *d_nn_v = nullptr, *d_nn_scalar = nullptr;
  void* d_nn_pack = nullptr;
  void* d_nn_tpack = nullptr;      // bf16 training packs (forward, biases, W^T tiles)
  float* d_nn_slabs = nullptr;     // per-64-sample gradient slabs of a training step
  size_t nn_slabs_cap = 0;
  float nn_lo[16] = {0}, nn_hi[16] = {0};
  bool nn_trained = false;
  uint64_t nn_version = 0;  // bumped whenever the weights or the input range change
  cvr_nn_report nn_report{};

  // volume gradient (float4 per voxel, brick order like the neighbour blocks)
  float4* d_grad = nullptr;
  uint64_t grad_built_for = ~0ull;

  // EBS directional shadow chain: host copy of the last table and its device copy
  std::vector<cvr::ShadowRow> ebs_rows, ebs_rows_dev;
  void* d_ebs_rows = nullptr;

  // object-space light cache (EBS, DOS): (occlusion, shadow) per cell, x fastest
  float2* d_lc = nullptr;
  size_t lc_cap = 0;
  int lc_res[3] = {0, 0, 0};
  bool lc_valid = false;

  // stats
  bool collect_stats = false;
  unsigned long long* d_counters = nullptr;
  unsigned int* d_nn_redo = nullptr;  // surrogate: tiles marked for the exact second pass
  size_t nn_redo_cap = 0;
  cvr_stats stats{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t ev_sat[3] = {nullptr, nullptr, nullptr};  // SAT build timing (stats on)

  // scratch output when rendering to a host buffer
  float* d_out = nullptr;
  size_t d_out_bytes = 0;
  // frame rendered at the multi-scale size before it is resampled to the screen
  float* d_ms_frame = nullptr;
  size_t d_ms_frame_bytes = 0;
};

namespace cvr {
cvr::VolGeom geom_of(const cvr_ctx* c);
cvr::RayFrame frame_of(const cvr_camera* cam, int W, int H);
float default_step(const cvr_ctx* c);
float volume_diagonal(const cvr_ctx* c);

// kernels (launchers return a CVR status)
int launch_rc1p(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st);
// Blinn-Phong constants of a render (builds the gradient when shading is on)
int phong_setup(cvr_ctx* c, int on, int gradient_type, float ks, float shininess, const float spec[3],
                const RayFrame& rf, const cvr_light* light, PhongK& ph, hipStream_t st);
// the SAT of the current volume and TF; with_blocks: also its corner blocks
int build_sat(cvr_ctx* c, hipStream_t st, bool with_blocks = false);
// volume gradient of `type` for Blinn-Phong shading
int build_gradient(cvr_ctx* c, int type, hipStream_t st);
// light cache: validates the resolution and sizes the buffer; the march over it
int lc_prepare(cvr_ctx* c, const int res[3]);
int launch_lc_render(cvr_ctx* c, const RayFrame& rf, int apply_occ, int apply_sdw, float ka, float kd, float step,
                     float et, const PhongK& ph, float* out, hipStream_t st);
int launch_ebs(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st);
int build_volume_oct(cvr_ctx* c, hipStream_t st);
int build_pyramid_oct(cvr_ctx* c, size_t total, int fmt, hipStream_t st);
int build_pyramid(cvr_ctx* c, int base_x, int base_y, int base_z, float sigma0, int nlev, int fmt, hipStream_t st);
PyrKey pyr_key_for(const cvr_ctx* c, int base_x, int base_y, int base_z, float sigma0, int nlev, int fmt);
// true when the pyramid and the cone table of P are built (a render allocates nothing)
bool cone_ready(const cvr_ctx* c, const cvr_cone_params& P);
int check_cone_params(const cvr_cone_params& P);
int prepare_cone(cvr_ctx* c, const cvr_cone_params& P, hipStream_t st);
int launch_cone(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st);
int cone_points(cvr_ctx* c, const cvr_cone_params& P, int n, const float* d_pos, const float* d_dirs,
                const float* d_lights, float* d_occ, float* d_sdw, float* d_feat, hipStream_t st);
int check_gt_params(const cvr_gt_params& P);
int prepare_gt(cvr_ctx* c, const cvr_gt_params& P);
int launch_gt(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st);
int check_dos_params(const cvr_dos_params& P);
int prepare_dos(cvr_ctx* c, const cvr_dos_params& P, hipStream_t st);
int launch_dos(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st);
int nn_alloc(cvr_ctx* c);
int nn_pack(cvr_ctx* c, hipStream_t st);
int nn_train(cvr_ctx* c, hipStream_t st);
int nn_gradient(cvr_ctx* c, const float* d_X, const float* d_T, int n, hipStream_t st);
int nn_eval(cvr_ctx* c, const float* d_X, int n, float* d_Y, int packed, hipStream_t st);
int launch_nn(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st);
int nn_param_count();
}  // namespace cvr
