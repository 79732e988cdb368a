// Technique instance state behind a cvr_handle.
#pragma once
#include "cvr_common.hpp"
#include "../../include/cvr.h"
#include <vector>

struct cvr_ctx {
  int technique = 0;
  int device = 0;

  // volume
  int nx = 0, ny = 0, nz = 0, dtype = 0;
  float sx = 1, sy = 1, sz = 1;
  void* d_vol = nullptr;

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

  // EBS summed-area table, two-float (hi, lo) per cell
  float2* d_sat = nullptr;
  float* d_ext_lut = nullptr;  // per raw value extinction (GetExtN), 256 or 65536 entries
  uint64_t sat_built_for = ~0ull;

  // stats
  bool collect_stats = false;
  unsigned long long* d_counters = nullptr;
  cvr_stats stats{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // scratch output when rendering to a host buffer
  float* d_out = nullptr;
  size_t d_out_bytes = 0;
};

namespace cvr {
cvr::VolGeom geom_of(const cvr_ctx* c);
cvr::RayFrame frame_of(const cvr_camera* cam, int W, int H);
float default_step(const cvr_ctx* c);
float volume_diagonal(const cvr_ctx* c);

// kernels (launchers return a CVR status)
int launch_rc1p(cvr_ctx* c, const RayFrame& rf, float* out, hipStream_t st);
int build_sat(cvr_ctx* c, hipStream_t st);
int launch_ebs(cvr_ctx* c, const RayFrame& rf, const cvr_light* light, float* out, hipStream_t st);
}  // namespace cvr
