// C ABI of the HIP ray casters: handle lifetime, inputs, parameters, dispatch.
// Declarations and the reference interface each entry point replaces: include/cvr.h.
#include "cvr_ctx.hpp"
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace cvr {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

VolGeom geom_of(const cvr_ctx* c) {
  VolGeom g{};
  g.nx = c->nx; g.ny = c->ny; g.nz = c->nz; g.dtype = c->dtype;
  const float s[3] = {c->sx, c->sy, c->sz};
  const int n[3] = {c->nx, c->ny, c->nz};
  for (int i = 0; i < 3; ++i) {
    g.scale[i] = s[i];
    // VolumeGridSize = resolution * voxel size (rc1prenderer.cpp:233-241), float
    g.G[i] = float(n[i]) * s[i];
    g.tex_mul[i] = float(double(n[i]) / double(g.G[i]));
  }
  g.vox = c->d_vol;
  g.oct = c->d_vol_oct;
  g.nbx = bricks(c->nx);
  g.nby = bricks(c->ny);
  for (int i = 0; i < 3; ++i) g.vmax[i] = float(n[i] - 1);
  return g;
}

// rc1prenderer.cpp:62-63: step = (0.5/sqrt(3)) * |voxel scale|, evaluated as the reference does.
float default_step(const cvr_ctx* c) {
  const double sx = c->sx, sy = c->sy, sz = c->sz;
  return float((0.5f / std::sqrt(3.0f)) * std::sqrt(sx * sx + sy * sy + sz * sz));
}

// StructuredGridVolume::GetDiagonal in the float form ebsrenderer.cpp:99-105 uses.
float volume_diagonal(const cvr_ctx* c) {
  const float w = float(c->nx) * c->sx, h = float(c->ny) * c->sy, d = float(c->nz) * c->sz;
  return std::sqrt(w * w + h * h + d * d);
}

// glm::lookAt basis in double (vis_utils/camera.cpp:281-284), rounded once to float.
RayFrame frame_of(const cvr_camera* cam, int W, int H) {
  RayFrame rf{};
  double e[3], c[3], u[3];
  for (int i = 0; i < 3; ++i) { e[i] = cam->eye[i]; c[i] = cam->center[i]; u[i] = cam->up[i]; }
  double f[3] = {c[0] - e[0], c[1] - e[1], c[2] - e[2]};
  double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (double& x : f) x /= fl;
  double s[3] = {f[1] * u[2] - f[2] * u[1], f[2] * u[0] - f[0] * u[2], f[0] * u[1] - f[1] * u[0]};
  double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (double& x : s) x /= sl;
  double uu[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
  for (int i = 0; i < 3; ++i) {
    rf.eye[i] = cam->eye[i];
    rf.s[i] = float(s[i]);
    rf.u[i] = float(uu[i]);
    rf.f[i] = float(f[i]);
  }
  rf.tanfov = float(std::tan((double(cam->fovy_deg) * (3.14159265358979323846 / 180.0)) / 2.0));
  rf.aspect = cam->aspect > 0 ? cam->aspect : float(W) / float(H);
  rf.W = W;
  rf.H = H;
  return rf;
}

static size_t params_size(int technique) {
  switch (technique) {
    case CVR_RC1P: return sizeof(cvr_rc1p_params);
    case CVR_EBS: return sizeof(cvr_ebs_params);
    case CVR_CONE_GT: return sizeof(cvr_gt_params);
    case CVR_DOS: return sizeof(cvr_dos_params);
    case CVR_CONE_TRACE: return sizeof(cvr_cone_params);
    case CVR_CONE_NN: return sizeof(cvr_nn_params);
    default: return 0;
  }
}

// Defaults of the cone-traced configuration: 8 occlusion cones of 30 degrees (1 axial,
// 7 on a ring at 60), one 10-degree shadow cone, 64 steps each, a 6-level pyramid at
// the volume's resolution with sigma0 = 1 (extcoefvolumegenerator.cpp:10-15 uses 1.0).
static void cone_defaults(cvr_cone_params* c) {
  c->step_size = 0.f;
  c->et_threshold = 0.99f;
  c->apply_occlusion = 1;
  c->apply_shadow = 1;
  c->n_levels = 6;
  c->base_res[0] = c->base_res[1] = c->base_res[2] = 0;
  c->sigma0 = 1.0f;
  c->n_steps = 64;
  c->n_occ_cones = 8;
  c->occ_half_deg = 30.f;
  c->ring_deg = 60.f;
  c->sdw_half_deg = 10.f;
  c->t0 = 1.0f;
  c->occ_max_dist = 0.f;
  c->sdw_max_dist = 0.f;
  c->r_sigma = 2.0f;
  c->ka = 0.5f;
  c->kd = 0.5f;
  c->density_precision = 0;
}

static void defaults(int technique, void* p) {
  if (technique == CVR_RC1P) {
    cvr_rc1p_params* r = static_cast<cvr_rc1p_params*>(p);
    r->step_size = 0.f;
    r->et_threshold = 0.99f;
    r->apply_gradient_shading = 0;  // rc1prenderer.cpp:22
    r->gradient_type = CVR_GRADIENT_SOBEL;
    r->ka = 0.5f;  // renderingparameters.cpp:23-26
    r->kd = 0.5f;
    r->ks = 0.8f;
    r->shininess = 30.0f;
    r->specular[0] = r->specular[1] = r->specular[2] = 1.0f;  // lightsourcelist.cpp:24
  } else if (technique == CVR_EBS) {
    cvr_ebs_params* e = static_cast<cvr_ebs_params*>(p);
    e->step_size = 0.f;
    e->et_threshold = 0.99f;
    e->apply_occlusion = 1;
    e->apply_shadow = 1;
    e->ao_shells = 15;
    e->ao_radius = 1.0f;
    e->sdw_angle_deg = 1.0f;
    e->sdw_interval = 2.0f;
    e->sdw_initial = 2.0f;
    e->sdw_weight = 1.0f;
    e->sdw_max_dist = 0.f;
    e->shadow_type = 0;
    e->ka = 0.5f;
    e->kd = 0.5f;
    e->light_cache = 0;  // ebsrenderer.cpp:46-47
    e->light_cache_res[0] = e->light_cache_res[1] = e->light_cache_res[2] = 32;
    e->apply_phong = 0;  // ebsrenderer.cpp:22
    e->gradient_type = CVR_GRADIENT_SOBEL;
    e->ks = 0.8f;  // renderingparameters.cpp:25-26
    e->shininess = 30.0f;
    e->specular[0] = e->specular[1] = e->specular[2] = 1.0f;  // lightsourcelist.cpp:24
    e->sat_layout = 0;  // corner blocks: 53.2 ms against 61.0 ms at 256^3 @1024^2
  } else if (technique == CVR_CONE_GT) {
    // crtgtrenderer.cpp:32-51 (both lighting terms on here: with both off the shading
    // divides by zero)
    cvr_gt_params* g = static_cast<cvr_gt_params*>(p);
    g->step_size = 0.f;
    g->et_threshold = 0.99f;
    g->apply_occlusion = 1;
    g->apply_shadow = 1;
    g->occ_rays = 1;
    g->occ_aperture_deg = 90.f;
    g->occ_max_dist = 0.f;
    g->sdw_rays = 1;
    g->sdw_aperture_deg = 1.f;
    g->sdw_max_dist = 0.f;
    g->shadow_type = 0;
    g->initial_gap = 1.0f;
    g->ray_step = 0.5f;
    g->ka = 0.5f;
    g->kd = 0.5f;
  } else if (technique == CVR_DOS) {
    // dosrcrenderer.cpp:29-70, conegaussiansampler.cpp:30-47, extcoefvolumegenerator.cpp:10-15
    cvr_dos_params* d = static_cast<cvr_dos_params*>(p);
    d->step_size = 0.f;
    d->et_threshold = 0.99f;
    d->apply_occlusion = 1;
    d->apply_shadow = 0;
    d->shadow_type = 0;
    d->spot_max_deg = 4.0f;
    d->n_levels = 0;
    d->base_res[0] = d->base_res[1] = d->base_res[2] = 128;
    d->sigma0 = 1.0f;
    d->occ = cvr_dos_cone{20.0f, 1, 0.35f, 3.0f, 0.f, 1.25f, 2.0f};
    d->sdw = cvr_dos_cone{0.5f, 0, 1.0f, 3.0f, 0.f, 1.25f, 2.0f};
    d->ka = 0.5f;
    d->kd = 0.5f;
    d->density_precision = 0;
    d->light_cache = 0;  // dosrcrenderer.cpp:64
    d->light_cache_res[0] = d->light_cache_res[1] = d->light_cache_res[2] = 32;
    d->apply_phong = 0;  // dosrcrenderer.cpp:33
    d->gradient_type = CVR_GRADIENT_SOBEL;
    d->ks = 0.8f;
    d->shininess = 30.0f;
    d->specular[0] = d->specular[1] = d->specular[2] = 1.0f;
  } else if (technique == CVR_CONE_TRACE) {
    cone_defaults(static_cast<cvr_cone_params*>(p));
  } else if (technique == CVR_CONE_NN) {
    cvr_nn_params* n = static_cast<cvr_nn_params*>(p);
    cone_defaults(&n->cone);
    n->cone.n_levels = 7;
    // the 16 taps read an fp16 pyramid (the reference's R16F storage, volvis_utils/utils.h:15,
    // extcoefvolumegenerator.cpp:113): 16-byte corner blocks, half the gathered bytes
    n->cone.density_precision = 1;
    n->n_train = 1 << 22;
    n->n_val = 1 << 16;
    n->epochs = 4;
    n->batch = 16384;
    n->lr = 2e-3f;
    n->rmse_gate = 0.05f;
    n->ood_margin = 0.02f;
    n->force_exact = 0;
    n->seed = 4;
    n->tile_w = 4;  // 4 x 16 pixels per wave: 4 % faster than 8 x 8 on the 512^3 surrogate frame
    n->shade_batch = 0.75f;
    n->span = 1;
  }
}

static void free_all(cvr_ctx* c) {
  if (c->d_vol) (void)hipFree(c->d_vol);
  if (c->d_tf_rgbt) (void)hipFree(c->d_tf_rgbt);
  if (c->d_tf_rgba) (void)hipFree(c->d_tf_rgba);
  if (c->d_sat) (void)hipFree(c->d_sat);
  if (c->d_sat_oct) (void)hipFree(c->d_sat_oct);
  if (c->d_ext_lut) (void)hipFree(c->d_ext_lut);
  if (c->d_counters) (void)hipFree(c->d_counters);
  if (c->d_out) (void)hipFree(c->d_out);
  if (c->d_ms_frame) (void)hipFree(c->d_ms_frame);
  if (c->d_pyr) (void)hipFree(c->d_pyr);
  if (c->d_pyr_oct) (void)hipFree(c->d_pyr_oct);
  if (c->d_nn_redo) (void)hipFree(c->d_nn_redo);
  if (c->d_vol_oct) (void)hipFree(c->d_vol_oct);
  if (c->d_cone_tab) (void)hipFree(c->d_cone_tab);
  if (c->d_dos_rows) (void)hipFree(c->d_dos_rows);
  if (c->d_gt_rays) (void)hipFree(c->d_gt_rays);
  if (c->d_lc) (void)hipFree(c->d_lc);
  if (c->d_grad) (void)hipFree(c->d_grad);
  if (c->d_ebs_rows) (void)hipFree(c->d_ebs_rows);
  for (void* p : {static_cast<void*>(c->d_nn_w), static_cast<void*>(c->d_nn_g), static_cast<void*>(c->d_nn_m),
                  static_cast<void*>(c->d_nn_v), static_cast<void*>(c->d_nn_scalar), c->d_nn_pack, c->d_nn_tpack,
                  static_cast<void*>(c->d_nn_slabs)})
    if (p) (void)hipFree(p);
  for (hipEvent_t e : c->ev_sat)
    if (e) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

}  // namespace cvr

using namespace cvr;

extern "C" {

const char* cvr_last_error(void) { return g_last_error.c_str(); }

const char* cvr_technique_name(int technique) {
  switch (technique) {
    case CVR_RC1P: return "1-Pass - Ray Casting";
    case CVR_EBS: return "1-Pass - Ray Casting - Extinction-based";
    case CVR_CONE_GT: return "1-Pass - Ray Casting - Cone Ground Truth (Steps)";
    case CVR_DOS: return "1-Pass - Ray Casting - Dir. Occlusion Shading";
    case CVR_CONE_TRACE: return "1-Pass - Ray Casting - Cone-traced Directional Occlusion";
    case CVR_CONE_NN: return "1-Pass - Ray Casting - Learned Cone Occlusion";
    default: return nullptr;
  }
}

int cvr_create(int technique, int device, cvr_handle* out) {
  if (!out) { set_error("cvr_create: out is null"); return CVR_ERR_INVALID; }
  *out = nullptr;
  if (params_size(technique) == 0) { set_error("cvr_create: unknown technique"); return CVR_ERR_UNSUPPORTED; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("cvr_create: no HIP device available");
    return CVR_ERR_HIP;
  }
  if (device < 0 || device >= ndev) { set_error("cvr_create: device index out of range"); return CVR_ERR_INVALID; }
  DeviceGuard dg(device);
  cvr_ctx* c = new cvr_ctx();
  c->technique = technique;
  c->device = device;
  defaults(CVR_RC1P, &c->rc1p);
  defaults(CVR_EBS, &c->ebs);
  defaults(CVR_CONE_GT, &c->gt);
  defaults(CVR_DOS, &c->dos);
  defaults(CVR_CONE_TRACE, &c->cone);
  defaults(CVR_CONE_NN, &c->nn);
  if (hipMalloc(&c->d_counters, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    free_all(c);
    delete c;
    set_error("cvr_create: HIP allocation failed");
    return CVR_ERR_HIP;
  }
  *out = c;
  return CVR_OK;
}

void cvr_destroy(cvr_handle h) {
  if (!h) return;
  DeviceGuard dg(h->device);
  (void)hipDeviceSynchronize();
  free_all(h);
  delete h;
}

int cvr_set_volume(cvr_handle h, int nx, int ny, int nz, int dtype, float sx, float sy, float sz,
                   const void* voxels, int is_device) {
  if (!h || !voxels) { set_error("cvr_set_volume: null handle or voxels"); return CVR_ERR_INVALID; }
  if (nx < 1 || ny < 1 || nz < 1 || (dtype != CVR_U8 && dtype != CVR_U16)) {
    set_error("cvr_set_volume: sizes must be >= 1 and dtype CVR_U8 or CVR_U16");
    return CVR_ERR_INVALID;
  }
  if (!(sx > 0) || !(sy > 0) || !(sz > 0)) { set_error("cvr_set_volume: voxel scale must be > 0"); return CVR_ERR_INVALID; }
  const size_t cells = size_t(nx) * ny * nz;
  if (cells > (size_t(1) << 34)) { set_error("cvr_set_volume: volume too large"); return CVR_ERR_INVALID; }
  DeviceGuard dg(h->device);
  const size_t bytes = cells * (dtype == CVR_U8 ? 1 : 2);
  const bool resize = !(h->d_vol && h->nx == nx && h->ny == ny && h->nz == nz && h->dtype == dtype);
  if (resize) {
    if (h->d_vol) CVR_HIP_CHECK(hipFree(h->d_vol));
    h->d_vol = nullptr;
    if (h->d_sat) CVR_HIP_CHECK(hipFree(h->d_sat));
    h->d_sat = nullptr;
    if (h->d_sat_oct) CVR_HIP_CHECK(hipFree(h->d_sat_oct));
    h->d_sat_oct = nullptr;
    if (h->d_grad) CVR_HIP_CHECK(hipFree(h->d_grad));
    h->d_grad = nullptr;
    if (hipMalloc(&h->d_vol, bytes) != hipSuccess) {
      h->d_vol = nullptr;
      set_error("cvr_set_volume: out of device memory");
      return CVR_ERR_NOMEM;
    }
  }
  CVR_HIP_CHECK(hipMemcpy(h->d_vol, voxels, bytes, is_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
  h->nx = nx; h->ny = ny; h->nz = nz; h->dtype = dtype;
  if (const int rc = build_volume_oct(h, nullptr)) return rc;
  h->sx = sx; h->sy = sy; h->sz = sz;
  h->vol_version++;
  h->nn_trained = false;
  return CVR_OK;
}

// TransferFunction1D::BuildLinear (transferfunction1d.cpp:319-358) with the float
// control-point colours of TransferControlPoint (transferfunction.cpp:8-24), then the
// RGBt / RGBA upload arrays (:58-118).
int cvr_set_transfer_function(cvr_handle h, int max_density, int extinction_input, int n_rgb,
                              const double* rgb_iso, int n_alpha, const double* alpha_iso) {
  if (!h) { set_error("cvr_set_transfer_function: null handle"); return CVR_ERR_INVALID; }
  if (max_density < 1 || max_density > (1 << 20)) {
    set_error("cvr_set_transfer_function: max_density out of range");
    return CVR_ERR_INVALID;
  }
  if (n_rgb < 0 || n_alpha < 0 || (n_rgb > 0 && !rgb_iso) || (n_alpha > 0 && !alpha_iso)) {
    set_error("cvr_set_transfer_function: bad control point arrays");
    return CVR_ERR_INVALID;
  }
  const int T = max_density + 1;
  std::vector<double> tab(size_t(T) * 4, 0.0);
  for (int i = 0; i + 1 < n_rgb; ++i) {
    const int i0 = int(rgb_iso[i * 4 + 3]), i1 = int(rgb_iso[(i + 1) * 4 + 3]);
    if (i0 == i1 || i0 < 0 || i1 < 0 || i0 >= T || i1 >= T) {
      set_error("cvr_set_transfer_function: rgb isovalues must be distinct and within [0, max_density]");
      return CVR_ERR_INVALID;
    }
    const float c0[3] = {float(rgb_iso[i * 4]), float(rgb_iso[i * 4 + 1]), float(rgb_iso[i * 4 + 2])};
    const float c1[3] = {float(rgb_iso[i * 4 + 4]), float(rgb_iso[i * 4 + 5]), float(rgb_iso[i * 4 + 6])};
    const double diff[3] = {double(c1[0] - c0[0]), double(c1[1] - c0[1]), double(c1[2] - c0[2])};
    for (int x = i0; x <= i1; ++x) {
      const double k = double(x - i0) / double(i1 - i0);
      for (int c = 0; c < 3; ++c) tab[size_t(x) * 4 + c] = double(c0[c]) + diff[c] * k;
    }
  }
  for (int i = 0; i + 1 < n_alpha; ++i) {
    const int i0 = int(alpha_iso[i * 2 + 1]), i1 = int(alpha_iso[(i + 1) * 2 + 1]);
    if (i0 == i1 || i0 < 0 || i1 < 0 || i0 >= T || i1 >= T) {
      set_error("cvr_set_transfer_function: alpha isovalues must be distinct and within [0, max_density]");
      return CVR_ERR_INVALID;
    }
    const float a0 = float(alpha_iso[i * 2]), a1 = float(alpha_iso[i * 2 + 2]);
    const double diff = double(a1 - a0);
    for (int x = i0; x <= i1; ++x) {
      const double k = double(x - i0) / double(i1 - i0);
      tab[size_t(x) * 4 + 3] = double(a0) + diff * k;
    }
  }
  std::vector<float4> rgbt(T), rgba(T);
  for (int i = 0; i < T; ++i) {
    const float r = float(tab[i * 4]), g = float(tab[i * 4 + 1]), b = float(tab[i * 4 + 2]);
    const float v4 = float(tab[i * 4 + 3]);
    const float ext = extinction_input ? v4 : float(std::log(1.0 / (1.0 - double(v4))));
    const float opc = extinction_input ? float(1.0 - std::exp(-double(v4))) : v4;
    rgbt[i] = make_float4(r, g, b, ext);
    rgba[i] = make_float4(r, g, b, opc);
  }
  DeviceGuard dg(h->device);
  if (h->tf_size != T) {
    if (h->d_tf_rgbt) CVR_HIP_CHECK(hipFree(h->d_tf_rgbt));
    if (h->d_tf_rgba) CVR_HIP_CHECK(hipFree(h->d_tf_rgba));
    h->d_tf_rgbt = h->d_tf_rgba = nullptr;
    if (hipMalloc(&h->d_tf_rgbt, T * sizeof(float4)) != hipSuccess ||
        hipMalloc(&h->d_tf_rgba, T * sizeof(float4)) != hipSuccess) {
      set_error("cvr_set_transfer_function: out of device memory");
      return CVR_ERR_NOMEM;
    }
  }
  CVR_HIP_CHECK(hipMemcpy(h->d_tf_rgbt, rgbt.data(), T * sizeof(float4), hipMemcpyHostToDevice));
  CVR_HIP_CHECK(hipMemcpy(h->d_tf_rgba, rgba.data(), T * sizeof(float4), hipMemcpyHostToDevice));
  h->tf_size = T;
  h->tf_max_density = max_density;
  h->tf_ext_input = extinction_input != 0;
  h->tf_tab.swap(tab);
  h->tf_version++;
  h->nn_trained = false;
  return CVR_OK;
}

int cvr_default_params(int technique, void* params, size_t bytes) {
  const size_t n = params_size(technique);
  if (n == 0) { set_error("cvr_default_params: unknown technique"); return CVR_ERR_UNSUPPORTED; }
  if (!params || bytes != n) { set_error("cvr_default_params: wrong params size"); return CVR_ERR_INVALID; }
  defaults(technique, params);
  return CVR_OK;
}

static void* live_params(cvr_ctx* c) {
  switch (c->technique) {
    case CVR_RC1P: return &c->rc1p;
    case CVR_EBS: return &c->ebs;
    case CVR_CONE_GT: return &c->gt;
    case CVR_DOS: return &c->dos;
    case CVR_CONE_TRACE: return &c->cone;
    case CVR_CONE_NN: return &c->nn;
    default: return nullptr;
  }
}

int cvr_set_params(cvr_handle h, const void* params, size_t bytes) {
  if (!h || !params) { set_error("cvr_set_params: null argument"); return CVR_ERR_INVALID; }
  if (bytes != params_size(h->technique)) { set_error("cvr_set_params: wrong params size for technique"); return CVR_ERR_INVALID; }
  if (h->technique == CVR_RC1P) {
    const cvr_rc1p_params* p = static_cast<const cvr_rc1p_params*>(params);
    if (!(p->et_threshold > 0) || !std::isfinite(p->step_size)) {
      set_error("cvr_set_params: et_threshold must be > 0");
      return CVR_ERR_INVALID;
    }
  }
  if (h->technique == CVR_CONE_GT) {
    const int rc = check_gt_params(*static_cast<const cvr_gt_params*>(params));
    if (rc) return rc;
  }
  if (h->technique == CVR_DOS) {
    const int rc = check_dos_params(*static_cast<const cvr_dos_params*>(params));
    if (rc) return rc;
  }
  {
    int gt = -1;
    if (h->technique == CVR_RC1P) gt = static_cast<const cvr_rc1p_params*>(params)->gradient_type;
    if (h->technique == CVR_EBS) gt = static_cast<const cvr_ebs_params*>(params)->gradient_type;
    if (h->technique == CVR_DOS) gt = static_cast<const cvr_dos_params*>(params)->gradient_type;
    if (gt != -1 && gt != CVR_GRADIENT_SOBEL && gt != CVR_GRADIENT_FINITE_DIFF) {
      set_error("cvr_set_params: gradient_type must be CVR_GRADIENT_SOBEL or CVR_GRADIENT_FINITE_DIFF");
      return CVR_ERR_INVALID;
    }
  }
  if (h->technique == CVR_EBS) {
    const int sl = static_cast<const cvr_ebs_params*>(params)->sat_layout;
    if (sl != 0 && sl != 1) {
      set_error("cvr_set_params: sat_layout must be 0 (corner blocks) or 1 (linear pairs)");
      return CVR_ERR_INVALID;
    }
  }
  if (h->technique == CVR_EBS || h->technique == CVR_DOS) {
    const int* r = h->technique == CVR_EBS ? static_cast<const cvr_ebs_params*>(params)->light_cache_res
                                           : static_cast<const cvr_dos_params*>(params)->light_cache_res;
    const int on = h->technique == CVR_EBS ? static_cast<const cvr_ebs_params*>(params)->light_cache
                                           : static_cast<const cvr_dos_params*>(params)->light_cache;
    for (int i = 0; i < 3 && on; ++i)
      if (r[i] < 1 || r[i] > 512) {
        set_error("cvr_set_params: light_cache_res must be in [1, 512] per axis");
        return CVR_ERR_INVALID;
      }
  }
  if (h->technique == CVR_CONE_TRACE || h->technique == CVR_CONE_NN) {
    const cvr_cone_params& cp = h->technique == CVR_CONE_TRACE ? *static_cast<const cvr_cone_params*>(params)
                                                                : static_cast<const cvr_nn_params*>(params)->cone;
    int rc = check_cone_params(cp);
    if (rc) return rc;
  }
  if (h->technique == CVR_CONE_NN) {
    const cvr_nn_params* q = static_cast<const cvr_nn_params*>(params);
    if ((q->tile_w != 4 && q->tile_w != 8 && q->tile_w != 16 && q->tile_w != 32) ||
        !(q->shade_batch >= 0.f && q->shade_batch <= 1.f) || (q->span != 1 && q->span != 2 && q->span != 4 && q->span != 8)) {
      set_error("cvr_set_params: tile_w must be 4, 8, 16 or 32, shade_batch in [0, 1], span 1, 2, 4 or 8");
      return CVR_ERR_INVALID;
    }
  }
  if (h->technique == CVR_CONE_NN) {
    // training inputs changed: the network is retrained on the next prepare/render
    // (the march step and early-termination threshold, the gate, the range margin and
    // force_exact apply at render time only)
    const cvr_nn_params* n = static_cast<const cvr_nn_params*>(params);
    const cvr_nn_params& o = h->nn;
    cvr_cone_params nc = n->cone, oc = o.cone;
    nc.step_size = oc.step_size = 0.f;
    nc.et_threshold = oc.et_threshold = 0.f;
    if (std::memcmp(&nc, &oc, sizeof(cvr_cone_params)) != 0 || n->n_train != o.n_train ||
        n->n_val != o.n_val || n->epochs != o.epochs || n->batch != o.batch || n->lr != o.lr || n->seed != o.seed)
      h->nn_trained = false;
  }
  std::memcpy(live_params(h), params, bytes);
  return CVR_OK;
}

int cvr_get_params(cvr_handle h, void* params, size_t bytes) {
  if (!h || !params) { set_error("cvr_get_params: null argument"); return CVR_ERR_INVALID; }
  if (bytes != params_size(h->technique)) { set_error("cvr_get_params: wrong params size for technique"); return CVR_ERR_INVALID; }
  std::memcpy(params, live_params(h), bytes);
  return CVR_OK;
}

static int check_inputs(cvr_ctx* c) {
  if (!c->d_vol) { set_error("no volume set (cvr_set_volume)"); return CVR_ERR_STATE; }
  if (!c->d_tf_rgbt) { set_error("no transfer function set (cvr_set_transfer_function)"); return CVR_ERR_STATE; }
  return CVR_OK;
}

int cvr_prepare(cvr_handle h, void* hip_stream) {
  if (!h) { set_error("cvr_prepare: null handle"); return CVR_ERR_INVALID; }
  int rc = check_inputs(h);
  if (rc) return rc;
  DeviceGuard dg(h->device);
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  if (h->technique == CVR_EBS) return build_sat(h, st, h->ebs.sat_layout == 0);
  if (h->technique == CVR_CONE_GT) return prepare_gt(h, h->gt);
  if (h->technique == CVR_DOS) return prepare_dos(h, h->dos, st);
  if (h->technique == CVR_CONE_TRACE) return prepare_cone(h, h->cone, st);
  if (h->technique == CVR_CONE_NN) {
    rc = prepare_cone(h, h->nn.cone, st);
    if (!rc && !h->nn_trained) rc = nn_train(h, st);
    return rc;
  }
  return CVR_OK;
}

}  // extern "C"

namespace cvr {

static int check_frame_args(const char* who, const cvr_camera* cam, int width, int height) {
  if (width < 1 || height < 1 || size_t(width) * height > (size_t(1) << 28)) {
    set_error(std::string(who) + ": image size out of range");
    return CVR_ERR_INVALID;
  }
  if (!(cam->fovy_deg > 0 && cam->fovy_deg < 180)) {
    set_error(std::string(who) + ": fovy must be in (0, 180)");
    return CVR_ERR_INVALID;
  }
  return CVR_OK;
}

// Grows a device buffer of the handle to at least `bytes`.
static int grow_buffer(float** p, size_t* cap, size_t bytes, const char* who) {
  if (*cap >= bytes) return CVR_OK;
  if (*p) CVR_HIP_CHECK(hipFree(*p));
  *p = nullptr;
  *cap = 0;
  if (hipMalloc(p, bytes) != hipSuccess) {
    *p = nullptr;
    set_error(std::string(who) + ": out of device memory");
    return CVR_ERR_NOMEM;
  }
  *cap = bytes;
  return CVR_OK;
}

// One frame of the technique into device memory `dst` (width x height RGBA floats),
// with the stats and the event timing of the handle.
static int render_device(cvr_ctx* h, const cvr_camera* cam, const cvr_light* light, int width, int height, float* dst,
                         hipStream_t st) {
  if (h->collect_stats) CVR_HIP_CHECK(hipMemsetAsync(h->d_counters, 0, 4 * sizeof(unsigned long long), st));
  RayFrame rf = frame_of(cam, width, height);
  CVR_HIP_CHECK(hipEventRecord(h->ev0, st));
  int rc;
  switch (h->technique) {
    case CVR_RC1P: rc = launch_rc1p(h, rf, light, dst, st); break;
    case CVR_EBS: rc = launch_ebs(h, rf, light, dst, st); break;
    case CVR_CONE_GT: rc = launch_gt(h, rf, light, dst, st); break;
    case CVR_DOS: rc = launch_dos(h, rf, light, dst, st); break;
    case CVR_CONE_TRACE: rc = launch_cone(h, rf, light, dst, st); break;
    case CVR_CONE_NN: rc = launch_nn(h, rf, light, dst, st); break;
    default: set_error("cvr_render: technique not available"); rc = CVR_ERR_UNSUPPORTED;
  }
  if (rc) return rc;
  CVR_HIP_CHECK(hipEventRecord(h->ev1, st));
  return CVR_OK;
}

// Reads the counters and the event time of the last render (waits for the stream).
static int finish_stats(cvr_ctx* h, hipStream_t st) {
  if (!h->collect_stats) return CVR_OK;
  unsigned long long cnt[4];
  CVR_HIP_CHECK(hipMemcpyAsync(cnt, h->d_counters, sizeof(cnt), hipMemcpyDeviceToHost, st));
  CVR_HIP_CHECK(hipStreamSynchronize(st));
  h->stats.samples = cnt[0];
  h->stats.shaded_samples = cnt[1];
  h->stats.fallback_samples = cnt[2];
  h->stats.box_sums = cnt[3];
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, h->ev0, h->ev1) == hipSuccess) h->stats.last_ms = ms;
  return CVR_OK;
}

}  // namespace cvr

extern "C" {

int cvr_render(cvr_handle h, const cvr_camera* cam, const cvr_light* light, int width, int height,
               float* out_rgba, int out_is_device, void* hip_stream) {
  if (!h || !cam || !out_rgba) { set_error("cvr_render: null argument"); return CVR_ERR_INVALID; }
  int rc = check_frame_args("cvr_render", cam, width, height);
  if (rc) return rc;
  if ((rc = check_inputs(h))) return rc;
  DeviceGuard dg(h->device);
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  const size_t bytes = size_t(width) * height * 4 * sizeof(float);
  float* dst = out_rgba;
  if (!out_is_device) {
    if ((rc = grow_buffer(&h->d_out, &h->d_out_bytes, bytes, "cvr_render"))) return rc;
    dst = h->d_out;
  }
  if ((rc = render_device(h, cam, light, width, height, dst, st))) return rc;
  if (!out_is_device) {
    CVR_HIP_CHECK(hipMemcpyAsync(out_rgba, dst, bytes, hipMemcpyDeviceToHost, st));
    CVR_HIP_CHECK(hipStreamSynchronize(st));
  }
  return finish_stats(h, st);
}

// BaseVolumeRenderer's multi-scale redraws (volrenderbase.h:28-33, 50-52): the technique
// renders a frame of the size Reshape gives the mode (volrenderbase.cpp:42-66: screen
// times the multipliers for multiple rays per pixel and down-scaling, screen divided by
// them for up-scaling), and RenderFrameToScreen's draw resamples it to the screen
// (DrawMultiSampleHigherResolutionMode, DrawHigherResolutionWithDownScale,
// DrawLowerResolutionWithUpScale, renderoutputframe.cpp:265-535).  The camera keeps the
// screen's aspect ratio, as the reference's camera follows the window.
int cvr_render_multiscale(cvr_handle h, const cvr_camera* cam, const cvr_light* light, int screen_w, int screen_h,
                          int mode, int mult_w, int mult_h, int kernel, float* out_rgba, int out_is_device,
                          void* hip_stream) {
  if (!h || !cam || !out_rgba) { set_error("cvr_render_multiscale: null argument"); return CVR_ERR_INVALID; }
  if (mode < CVR_SINGLE_RAY_PER_PIXEL || mode > CVR_UP_SCALING_RENDER) {
    set_error("cvr_render_multiscale: unknown multi-scaling mode");
    return CVR_ERR_INVALID;
  }
  if (mode == CVR_SINGLE_RAY_PER_PIXEL)
    return cvr_render(h, cam, light, screen_w, screen_h, out_rgba, out_is_device, hip_stream);
  if (mult_w < 1 || mult_h < 1 || kernel < CVR_FILTER_BOX || kernel > CVR_FILTER_CARDINAL_OMOMS3) {
    set_error("cvr_render_multiscale: multipliers must be >= 1 and the kernel a CVR_FILTER_* value");
    return CVR_ERR_INVALID;
  }
  int rc = check_frame_args("cvr_render_multiscale", cam, screen_w, screen_h);
  if (rc) return rc;
  if ((rc = check_inputs(h))) return rc;
  const int sgn = mode == CVR_UP_SCALING_RENDER ? -1 : 1;
  int fw = 0, fh = 0;
  if ((rc = cvr_multiscale_size(screen_w, screen_h, sgn * mult_w, sgn * mult_h, &fw, &fh))) return rc;
  if ((rc = check_frame_args("cvr_render_multiscale", cam, fw, fh))) return rc;
  DeviceGuard dg(h->device);
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  const size_t fbytes = size_t(fw) * fh * 4 * sizeof(float), sbytes = size_t(screen_w) * screen_h * 4 * sizeof(float);
  if ((rc = grow_buffer(&h->d_ms_frame, &h->d_ms_frame_bytes, fbytes, "cvr_render_multiscale"))) return rc;
  float* dst = out_rgba;
  if (!out_is_device) {
    if ((rc = grow_buffer(&h->d_out, &h->d_out_bytes, sbytes, "cvr_render_multiscale"))) return rc;
    dst = h->d_out;
  }
  cvr_camera fc = *cam;
  if (!(fc.aspect > 0)) fc.aspect = float(screen_w) / float(screen_h);
  if ((rc = render_device(h, &fc, light, fw, fh, h->d_ms_frame, st))) return rc;
  const int smode = mode == CVR_MULTIPLE_RAYS_PER_PIXEL ? CVR_SCALE_MULTISAMPLE
                    : mode == CVR_DOWN_SCALING_RENDER   ? CVR_SCALE_DOWNSCALE
                                                        : CVR_SCALE_UPSCALE;
  // the cardinal up-scaling prefilters the rendered frame in place, as the reference does
  if ((rc = cvr_resample_frame(h->device, h->d_ms_frame, fw, fh, dst, screen_w, screen_h, smode, kernel,
                               h->d_ms_frame, hip_stream)))
    return rc;
  if (!out_is_device) {
    CVR_HIP_CHECK(hipMemcpyAsync(out_rgba, dst, sbytes, hipMemcpyDeviceToHost, st));
    CVR_HIP_CHECK(hipStreamSynchronize(st));
  }
  return finish_stats(h, st);
}

int cvr_get_sat(cvr_handle h, double* out_host, size_t count) {
  if (!h || !out_host) { set_error("cvr_get_sat: null argument"); return CVR_ERR_INVALID; }
  if (h->technique != CVR_EBS) { set_error("cvr_get_sat: technique has no SAT"); return CVR_ERR_UNSUPPORTED; }
  int rc = check_inputs(h);
  if (rc) return rc;
  const size_t cells = size_t(h->nx + 2) * (h->ny + 2) * (h->nz + 2);
  if (count != cells) { set_error("cvr_get_sat: count must be (nx+2)(ny+2)(nz+2)"); return CVR_ERR_INVALID; }
  DeviceGuard dg(h->device);
  rc = build_sat(h, nullptr);
  if (rc) return rc;
  std::vector<float2> tmp(cells);
  CVR_HIP_CHECK(hipMemcpy(tmp.data(), h->d_sat, cells * sizeof(float2), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < cells; ++i) out_host[i] = double(tmp[i].x) + double(tmp[i].y);
  return CVR_OK;
}

int cvr_get_gradient(cvr_handle h, int type, float* out_host, size_t count) {
  if (!h || !out_host) { set_error("cvr_get_gradient: null argument"); return CVR_ERR_INVALID; }
  if (type != CVR_GRADIENT_SOBEL && type != CVR_GRADIENT_FINITE_DIFF) {
    set_error("cvr_get_gradient: unknown gradient type");
    return CVR_ERR_INVALID;
  }
  int rc = check_inputs(h);
  if (rc) return rc;
  const size_t n = size_t(h->nx) * h->ny * h->nz;
  if (count < 3 * n) { set_error("cvr_get_gradient: output too small"); return CVR_ERR_INVALID; }
  DeviceGuard dg(h->device);
  if ((rc = build_gradient(h, type, nullptr))) return rc;
  const int nbx = bricks(h->nx), nby = bricks(h->ny);
  std::vector<float4> tmp(brick_cells(h->nx, h->ny, h->nz));
  CVR_HIP_CHECK(hipMemcpy(tmp.data(), h->d_grad, tmp.size() * sizeof(float4), hipMemcpyDeviceToHost));
  size_t o = 0;
  for (int z = 0; z < h->nz; ++z)
    for (int y = 0; y < h->ny; ++y)
      for (int x = 0; x < h->nx; ++x, o += 3) {
        const float4 g = tmp[brick_index(x, y, z, nbx, nby)];
        out_host[o] = g.x;
        out_host[o + 1] = g.y;
        out_host[o + 2] = g.z;
      }
  return CVR_OK;
}

int cvr_get_light_cache(cvr_handle h, float* out_host, size_t count, int* res) {
  if (!h || !res) { set_error("cvr_get_light_cache: null argument"); return CVR_ERR_INVALID; }
  if (!h->lc_valid) { set_error("cvr_get_light_cache: no light cache rendered yet"); return CVR_ERR_STATE; }
  for (int i = 0; i < 3; ++i) res[i] = h->lc_res[i];
  const size_t n = size_t(h->lc_res[0]) * h->lc_res[1] * h->lc_res[2] * 2;
  if (!out_host) return CVR_OK;
  if (count < n) { set_error("cvr_get_light_cache: output too small"); return CVR_ERR_INVALID; }
  DeviceGuard dg(h->device);
  CVR_HIP_CHECK(hipDeviceSynchronize());
  CVR_HIP_CHECK(hipMemcpy(out_host, h->d_lc, n * sizeof(float), hipMemcpyDeviceToHost));
  return CVR_OK;
}

static const cvr_cone_params* cone_of(cvr_ctx* h) {
  if (h->technique == CVR_CONE_TRACE) return &h->cone;
  if (h->technique == CVR_CONE_NN) return &h->nn.cone;
  return nullptr;
}

// Builds the pyramid of a pyramid-based technique.
static int prepare_pyramid(cvr_ctx* h, const char* who) {
  const cvr_cone_params* cp = cone_of(h);
  if (!cp && h->technique != CVR_DOS) {
    set_error(std::string(who) + ": technique has no pyramid");
    return CVR_ERR_UNSUPPORTED;
  }
  int rc = check_inputs(h);
  if (rc) return rc;
  return cp ? prepare_cone(h, *cp, nullptr) : prepare_dos(h, h->dos, nullptr);
}

int cvr_pyramid_info(cvr_handle h, int* levels, int* dims, int max_levels) {
  if (!h || !levels) { set_error("cvr_pyramid_info: null argument"); return CVR_ERR_INVALID; }
  DeviceGuard dg(h->device);
  int rc = prepare_pyramid(h, "cvr_pyramid_info");
  if (rc) return rc;
  *levels = h->pyr.levels;
  if (dims)
    for (int l = 0; l < std::min(max_levels, h->pyr.levels); ++l)
      for (int k = 0; k < 3; ++k) dims[3 * l + k] = h->pyr.dims[l][k];
  return CVR_OK;
}

int cvr_get_pyramid_level(cvr_handle h, int level, float* out_host, size_t count) {
  if (!h || !out_host) { set_error("cvr_get_pyramid_level: null argument"); return CVR_ERR_INVALID; }
  DeviceGuard dg(h->device);
  int rc = prepare_pyramid(h, "cvr_get_pyramid_level");
  if (rc) return rc;
  if (level < 0 || level >= h->pyr.levels) { set_error("cvr_get_pyramid_level: level out of range"); return CVR_ERR_INVALID; }
  const size_t n = size_t(h->pyr.dims[level][0]) * h->pyr.dims[level][1] * h->pyr.dims[level][2];
  if (count != n) { set_error("cvr_get_pyramid_level: count must be the level's voxel count"); return CVR_ERR_INVALID; }
  CVR_HIP_CHECK(hipMemcpy(out_host, h->d_pyr + h->pyr.off[level], n * sizeof(float), hipMemcpyDeviceToHost));
  return CVR_OK;
}

// Temporary device copies for host-array entry points.
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
};

int cvr_cone_eval(cvr_handle h, int n, const float* pos, const float* dirs, const float* lights, float* occ,
                  float* sdw, float* feat) {
  if (!h || !pos || !dirs || !lights || n < 0) { set_error("cvr_cone_eval: null argument"); return CVR_ERR_INVALID; }
  const cvr_cone_params* cp = cone_of(h);
  if (!cp) { set_error("cvr_cone_eval: technique has no cone tracer"); return CVR_ERR_UNSUPPORTED; }
  int rc = check_inputs(h);
  if (rc || n == 0) return rc;
  DeviceGuard dg(h->device);
  DevBuf in, o;
  const size_t n3 = size_t(n) * 3;
  CVR_HIP_CHECK(hipMalloc(&in.p, 3 * n3 * sizeof(float)));
  CVR_HIP_CHECK(hipMalloc(&o.p, size_t(n) * 18 * sizeof(float)));
  float* dp = static_cast<float*>(in.p);
  float* dout = static_cast<float*>(o.p);
  CVR_HIP_CHECK(hipMemcpy(dp, pos, n3 * 4, hipMemcpyHostToDevice));
  CVR_HIP_CHECK(hipMemcpy(dp + n3, dirs, n3 * 4, hipMemcpyHostToDevice));
  CVR_HIP_CHECK(hipMemcpy(dp + 2 * n3, lights, n3 * 4, hipMemcpyHostToDevice));
  rc = cone_points(h, *cp, n, dp, dp + n3, dp + 2 * n3, dout, dout + n, dout + 2 * size_t(n), nullptr);
  if (rc) return rc;
  CVR_HIP_CHECK(hipDeviceSynchronize());
  if (occ) CVR_HIP_CHECK(hipMemcpy(occ, dout, size_t(n) * 4, hipMemcpyDeviceToHost));
  if (sdw) CVR_HIP_CHECK(hipMemcpy(sdw, dout + n, size_t(n) * 4, hipMemcpyDeviceToHost));
  if (feat) CVR_HIP_CHECK(hipMemcpy(feat, dout + 2 * size_t(n), size_t(n) * 16 * 4, hipMemcpyDeviceToHost));
  return CVR_OK;
}

int cvr_nn_param_count(void) { return nn_param_count(); }

static int nn_check(cvr_ctx* h) {
  if (!h) { set_error("surrogate: null handle"); return CVR_ERR_INVALID; }
  if (h->technique != CVR_CONE_NN) { set_error("surrogate: technique is not CVR_CONE_NN"); return CVR_ERR_UNSUPPORTED; }
  return CVR_OK;
}

int cvr_nn_train(cvr_handle h, void* hip_stream) {
  int rc = nn_check(h);
  if (rc) return rc;
  rc = check_inputs(h);
  if (rc) return rc;
  DeviceGuard dg(h->device);
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  rc = prepare_cone(h, h->nn.cone, st);
  if (!rc) rc = nn_train(h, st);
  if (!rc) CVR_HIP_CHECK(hipStreamSynchronize(st));
  return rc;
}

int cvr_nn_get_weights(cvr_handle h, float* out, size_t count) {
  int rc = nn_check(h);
  if (rc) return rc;
  if (!out || count != size_t(nn_param_count())) { set_error("cvr_nn_get_weights: count must be cvr_nn_param_count()"); return CVR_ERR_INVALID; }
  if (!h->d_nn_w) { set_error("cvr_nn_get_weights: no network yet"); return CVR_ERR_STATE; }
  DeviceGuard dg(h->device);
  CVR_HIP_CHECK(hipDeviceSynchronize());
  CVR_HIP_CHECK(hipMemcpy(out, h->d_nn_w, count * 4, hipMemcpyDeviceToHost));
  return CVR_OK;
}

int cvr_nn_set_weights(cvr_handle h, const float* in, size_t count) {
  int rc = nn_check(h);
  if (rc) return rc;
  if (!in || count != size_t(nn_param_count())) { set_error("cvr_nn_set_weights: count must be cvr_nn_param_count()"); return CVR_ERR_INVALID; }
  DeviceGuard dg(h->device);
  rc = nn_alloc(h);
  if (rc) return rc;
  CVR_HIP_CHECK(hipMemcpy(h->d_nn_w, in, count * 4, hipMemcpyHostToDevice));
  rc = nn_pack(h, nullptr);
  if (rc) return rc;
  CVR_HIP_CHECK(hipDeviceSynchronize());
  // no range information: every input is in range, and the gate is open
  for (int i = 0; i < 16; ++i) { h->nn_lo[i] = -1e30f; h->nn_hi[i] = 1e30f; }
  h->nn_report = cvr_nn_report{};
  h->nn_trained = true;
  h->nn_version++;
  return CVR_OK;
}

int cvr_nn_eval(cvr_handle h, int n, const float* X, float* Y, int path) {
  int rc = nn_check(h);
  if (rc) return rc;
  if (!X || !Y || n < 0) { set_error("cvr_nn_eval: null argument"); return CVR_ERR_INVALID; }
  if (!h->d_nn_w) { set_error("cvr_nn_eval: no network yet"); return CVR_ERR_STATE; }
  if (n == 0) return CVR_OK;
  DeviceGuard dg(h->device);
  DevBuf dx, dy;
  CVR_HIP_CHECK(hipMalloc(&dx.p, size_t(n) * 16 * 4));
  CVR_HIP_CHECK(hipMalloc(&dy.p, size_t(n) * 2 * 4));
  CVR_HIP_CHECK(hipMemcpy(dx.p, X, size_t(n) * 16 * 4, hipMemcpyHostToDevice));
  rc = nn_eval(h, static_cast<float*>(dx.p), n, static_cast<float*>(dy.p), path, nullptr);
  if (rc) return rc;
  CVR_HIP_CHECK(hipDeviceSynchronize());
  CVR_HIP_CHECK(hipMemcpy(Y, dy.p, size_t(n) * 2 * 4, hipMemcpyDeviceToHost));
  return CVR_OK;
}

int cvr_nn_gradient(cvr_handle h, int n, const float* X, const float* T, float* grad, float* loss) {
  int rc = nn_check(h);
  if (rc) return rc;
  if (!X || !T || !grad || n < 1) { set_error("cvr_nn_gradient: null argument"); return CVR_ERR_INVALID; }
  if (!h->d_nn_w) { set_error("cvr_nn_gradient: no network yet"); return CVR_ERR_STATE; }
  DeviceGuard dg(h->device);
  DevBuf dx, dt;
  CVR_HIP_CHECK(hipMalloc(&dx.p, size_t(n) * 16 * 4));
  CVR_HIP_CHECK(hipMalloc(&dt.p, size_t(n) * 2 * 4));
  CVR_HIP_CHECK(hipMemcpy(dx.p, X, size_t(n) * 16 * 4, hipMemcpyHostToDevice));
  CVR_HIP_CHECK(hipMemcpy(dt.p, T, size_t(n) * 2 * 4, hipMemcpyHostToDevice));
  rc = nn_gradient(h, static_cast<float*>(dx.p), static_cast<float*>(dt.p), n, nullptr);
  if (rc) return rc;
  CVR_HIP_CHECK(hipDeviceSynchronize());
  CVR_HIP_CHECK(hipMemcpy(grad, h->d_nn_g, size_t(nn_param_count()) * 4, hipMemcpyDeviceToHost));
  if (loss) CVR_HIP_CHECK(hipMemcpy(loss, h->d_nn_scalar, 4, hipMemcpyDeviceToHost));
  return CVR_OK;
}

int cvr_nn_get_report(cvr_handle h, cvr_nn_report* out) {
  int rc = nn_check(h);
  if (rc) return rc;
  if (!out) { set_error("cvr_nn_get_report: null argument"); return CVR_ERR_INVALID; }
  *out = h->nn_report;
  return CVR_OK;
}

// Checkpoint layout: "CVRNNCK1" | u32 version | u32 nparams | w | m | v (fp32 each) |
// lo[16] | hi[16] | cvr_nn_report.
static const char kCkMagic[8] = {'C', 'V', 'R', 'N', 'N', 'C', 'K', '1'};
size_t cvr_nn_checkpoint_size(void) {
  return 8 + 4 + 4 + 3 * size_t(nn_param_count()) * 4 + 32 * 4 + sizeof(cvr_nn_report);
}

int cvr_nn_save(cvr_handle h, void* out, size_t bytes) {
  int rc = nn_check(h);
  if (rc) return rc;
  if (!out || bytes != cvr_nn_checkpoint_size()) { set_error("cvr_nn_save: buffer must be cvr_nn_checkpoint_size() bytes"); return CVR_ERR_INVALID; }
  if (!h->d_nn_w || !h->nn_trained) { set_error("cvr_nn_save: no trained network"); return CVR_ERR_STATE; }
  DeviceGuard dg(h->device);
  CVR_HIP_CHECK(hipDeviceSynchronize());
  char* p = static_cast<char*>(out);
  const uint32_t ver = 1, np = uint32_t(nn_param_count());
  std::memcpy(p, kCkMagic, 8); p += 8;
  std::memcpy(p, &ver, 4); p += 4;
  std::memcpy(p, &np, 4); p += 4;
  for (float* d : {h->d_nn_w, h->d_nn_m, h->d_nn_v}) {
    CVR_HIP_CHECK(hipMemcpy(p, d, size_t(np) * 4, hipMemcpyDeviceToHost));
    p += size_t(np) * 4;
  }
  std::memcpy(p, h->nn_lo, 64); p += 64;
  std::memcpy(p, h->nn_hi, 64); p += 64;
  std::memcpy(p, &h->nn_report, sizeof(cvr_nn_report));
  return CVR_OK;
}

int cvr_nn_load(cvr_handle h, const void* in, size_t bytes) {
  int rc = nn_check(h);
  if (rc) return rc;
  if (!in || bytes != cvr_nn_checkpoint_size()) { set_error("cvr_nn_load: size mismatch"); return CVR_ERR_INVALID; }
  const char* p = static_cast<const char*>(in);
  uint32_t ver = 0, np = 0;
  std::memcpy(&ver, p + 8, 4);
  std::memcpy(&np, p + 12, 4);
  if (std::memcmp(p, kCkMagic, 8) != 0 || ver != 1 || np != uint32_t(nn_param_count())) {
    set_error("cvr_nn_load: not a version-1 surrogate checkpoint");
    return CVR_ERR_INVALID;
  }
  p += 16;
  DeviceGuard dg(h->device);
  rc = nn_alloc(h);
  if (rc) return rc;
  for (float* d : {h->d_nn_w, h->d_nn_m, h->d_nn_v}) {
    CVR_HIP_CHECK(hipMemcpy(d, p, size_t(np) * 4, hipMemcpyHostToDevice));
    p += size_t(np) * 4;
  }
  std::memcpy(h->nn_lo, p, 64); p += 64;
  std::memcpy(h->nn_hi, p, 64); p += 64;
  std::memcpy(&h->nn_report, p, sizeof(cvr_nn_report));
  rc = nn_pack(h, nullptr);
  if (rc) return rc;
  CVR_HIP_CHECK(hipDeviceSynchronize());
  h->nn_trained = true;
  h->nn_version++;
  return CVR_OK;
}

int cvr_dos_sections(cvr_handle h, int which, float* rows, int max_rows, int* n_sections, int* counts, float* axes,
                     float* adj7) {
  if (!h) { set_error("cvr_dos_sections: null handle"); return CVR_ERR_INVALID; }
  if (h->technique != CVR_DOS) { set_error("cvr_dos_sections: technique has no section tables"); return CVR_ERR_UNSUPPORTED; }
  if (which != 0 && which != 1) { set_error("cvr_dos_sections: which must be 0 (occlusion) or 1 (shadow)"); return CVR_ERR_INVALID; }
  if (rows && max_rows < 0) { set_error("cvr_dos_sections: max_rows < 0"); return CVR_ERR_INVALID; }
  int rc = check_inputs(h);
  if (rc) return rc;
  DeviceGuard dg(h->device);
  rc = prepare_dos(h, h->dos, nullptr);
  if (rc) return rc;
  const DosTable& T = which == 0 ? h->dos_occ : h->dos_sdw;
  const int n = int(T.rows.size());
  if (n_sections) *n_sections = n;
  if (rows)
    for (int i = 0; i < std::min(n, max_rows); ++i) {
      rows[4 * i] = T.rows[i].x;
      rows[4 * i + 1] = T.rows[i].y;
      rows[4 * i + 2] = T.rows[i].z;
      rows[4 * i + 3] = T.rows[i].w;
    }
  if (counts)
    for (int k = 0; k < 3; ++k) counts[k] = T.n[k];
  if (axes) std::memcpy(axes, T.axes, sizeof(T.axes));
  if (adj7) *adj7 = T.adj7;
  return CVR_OK;
}

int cvr_gt_rays(cvr_handle h, int which, float* out, int max_rays, int* n) {
  if (!h) { set_error("cvr_gt_rays: null handle"); return CVR_ERR_INVALID; }
  if (h->technique != CVR_CONE_GT) { set_error("cvr_gt_rays: technique has no ray tables"); return CVR_ERR_UNSUPPORTED; }
  if (which != 0 && which != 1) { set_error("cvr_gt_rays: which must be 0 (occlusion) or 1 (shadow)"); return CVR_ERR_INVALID; }
  DeviceGuard dg(h->device);
  int rc = prepare_gt(h, h->gt);
  if (rc) return rc;
  const std::vector<float>& R = which == 0 ? h->gt_occ : h->gt_sdw;
  const int m = int(R.size() / 3);
  if (n) *n = m;
  if (out)
    for (int i = 0; i < std::min(m, max_rays) * 3; ++i) out[i] = R[i];
  return CVR_OK;
}

int cvr_get_stats(cvr_handle h, cvr_stats* out) {
  if (!h || !out) { set_error("cvr_get_stats: null argument"); return CVR_ERR_INVALID; }
  *out = h->stats;
  return CVR_OK;
}

int cvr_set_collect_stats(cvr_handle h, int enable) {
  if (!h) { set_error("cvr_set_collect_stats: null handle"); return CVR_ERR_INVALID; }
  h->collect_stats = enable != 0;
  return CVR_OK;
}

}  // extern "C"
