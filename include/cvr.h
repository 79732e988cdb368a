/*
 * cvr.h — C ABI of the MI355X-native single-pass structured volume ray casters.
 *
 * One handle = one technique instance that owns all of its device memory, the
 * role a BaseVolumeRenderer subclass plays in the reference
 * (cppvolrend/volrenderbase.h:25-98).  Plain pointers and sizes only; status 0
 * is success, negative values are errors (cvr_last_error() has the message).
 * The ABI never exits the process.  A handle is not thread-safe: use one
 * handle per host thread.  Inputs are copied; caller buffers stay the
 * caller's.
 *
 * Coordinates follow the reference: the volume box is [-G/2, G/2] with
 * G = resolution * voxel scale, the image is premultiplied RGBA with row 0 at
 * the bottom (GL image convention), rays per pixel follow
 * cppvolrend/structured/rc1pass/ray_marching_1p.comp:158-175.
 */
#ifndef CVR_H
#define CVR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CVR_OK 0
#define CVR_ERR_INVALID (-1)     /* bad argument: null pointer, size, out-of-range parameter */
#define CVR_ERR_HIP (-2)         /* a HIP runtime call failed */
#define CVR_ERR_STATE (-3)       /* call order: e.g. render before a volume / transfer function is set */
#define CVR_ERR_UNSUPPORTED (-4) /* technique or option not available */
#define CVR_ERR_NOMEM (-5)       /* device allocation failed */

/* Techniques.  Each replaces one reference renderer class:
 *   CVR_RC1P        RayCasting1Pass            cppvolrend/structured/rc1pass/rc1prenderer.h:32-78
 *   CVR_EBS         RC1PExtinctionBasedShading cppvolrend/structured/rc1pextbsd/ebsrenderer.h:45-119
 *   CVR_CONE_GT     RC1PConeLightGroundTruthSteps (ray-sampled cone ground truth)
 *                   cppvolrend/structured/rc1pcrtgt/crtgtrenderer.h:50-120
 *   CVR_DOS         RC1PConeTracingDirOcclusionShading (section tables of
 *                   ConeGaussianSampler integrated over the Gaussian extinction
 *                   pyramid) cppvolrend/structured/rc1pdosct/dosrcrenderer.h:38-121
 *   CVR_CONE_TRACE  cone-traced directional occlusion + shadows over the Gaussian
 *                   extinction pyramid of RC1PConeTracingDirOcclusionShading
 *                   (cppvolrend/structured/rc1pdosct/dosrcrenderer.h:38-121 and
 *                   extcoefvolumegenerator.cpp:230-408), with 1 axial + ring
 *                   occlusion cones and one shadow cone marched in fixed steps
 *   CVR_CONE_NN     the same renderer with the per-sample cone integration replaced
 *                   by a learned bf16 surrogate; no reference counterpart
 */
enum cvr_technique { CVR_RC1P = 0, CVR_EBS = 1, CVR_CONE_GT = 2, CVR_DOS = 3, CVR_CONE_TRACE = 4, CVR_CONE_NN = 5 };

enum cvr_dtype { CVR_U8 = 0, CVR_U16 = 1 };

typedef struct cvr_ctx* cvr_handle;

/* Camera (vis_utils/camera.cpp:17-34, 281-310): glm::lookAt(eye, center, up),
 * vertical field of view in degrees, aspect <= 0 means width/height. */
typedef struct cvr_camera {
  float eye[3];
  float center[3];
  float up[3];
  float fovy_deg;
  float aspect;
} cvr_camera;

/* Light (volvis_utils/renderingparameters.cpp:147-167): point-light position and
 * the light camera axes; `forward` points from the scene towards the light. */
typedef struct cvr_light {
  float position[3];
  float forward[3];
  float up[3];
  float right[3];
} cvr_light;

/* RayCasting1Pass parameters (rc1prenderer.cpp:18-23, 62-63). */
/* Volume gradient for Blinn-Phong shading (DataManager::GenerateStructuredGradientTexture,
 * datamanager.cpp:332-352): Sobel-Feldman (utils.cpp:287-350 and the compute-shader
 * form _gradient_shading/sobelfeldman_generator.comp, unnormalised, 0 outside the
 * volume) or central finite differences (utils.cpp:146-284 with its defaults: step 1,
 * no filter, normalised, 0 where the difference vanishes). */
enum { CVR_GRADIENT_SOBEL = 0, CVR_GRADIENT_FINITE_DIFF = 1 };

typedef struct cvr_rc1p_params {
  float step_size;    /* <= 0: reference default 0.5/sqrt(3)*|voxel scale| */
  float et_threshold; /* early termination, ray_marching_1p.comp:238 (0.99) */
  /* ShadeBlinnPhong (ray_marching_1p.comp:48-81): rgb*(ka + kd*max(0, n.l)) +
   * specular*ks*max(0, h.n)^shininess where the gradient is nonzero; the light is
   * cvr_light.position (RenderingParameters, renderingparameters.cpp:17-31: 0.5, 0.5,
   * 0.8, 30; specular (1, 1, 1), lightsourcelist.cpp:24).  Off by default
   * (rc1prenderer.cpp:22). */
  int apply_gradient_shading;
  int gradient_type;  /* CVR_GRADIENT_SOBEL or CVR_GRADIENT_FINITE_DIFF */
  float ka, kd, ks, shininess;
  float specular[3];
} cvr_rc1p_params;

/* RC1PExtinctionBasedShading parameters (ebsrenderer.cpp:19-53, 99-105, 125-247). */
typedef struct cvr_ebs_params {
  float step_size;       /* <= 0: reference default */
  float et_threshold;    /* 0.99 */
  int apply_occlusion;   /* ApplyOcclusion */
  int apply_shadow;      /* ApplyShadow */
  int ao_shells;         /* AmbOccShells, 15 */
  float ao_radius;       /* AmbOccRadius, 1.0 (voxels) */
  float sdw_angle_deg;   /* DirSdwConeAngle in degrees (half angle), 1.0 */
  float sdw_interval;    /* DirSdwSampleInterval, 2.0 */
  float sdw_initial;     /* DirSdwInitialStep, 2.0 */
  float sdw_weight;      /* DirSdwUserInterfaceWeight, 1.0 */
  float sdw_max_dist;    /* DirSdwConeMaxDistance; <= 0: 0.75 * volume diagonal */
  int shadow_type;       /* TypeOfShadow: 0 point light (position), 1 directional (forward) */
  float ka, kd;          /* Kambient, Kdiffuse: 0.5, 0.5 */
  /* PreIlluminationStructuredVolume (preillumination.cpp:7-80, ebsrenderer.cpp:46-47,
   * 127-134): nonzero renders through an object-space light cache of
   * light_cache_res cells (occlusion, shadow per cell, rebuilt every render,
   * lightcachecomputation.comp:440-468) read trilinearly by the march
   * (_common_shaders/obj_ray_marching.comp:223-333).  Off, 32^3. */
  int light_cache;
  int light_cache_res[3];
  /* ApplyPhongShading (ebs_ray_bbox_marching.comp:523-546, ebsrenderer.cpp:22, 221):
   * rgb <- (rgb*Iocc*ka + Isdw*rgb*kd*max(0, n.l)) / (ka+kd) + Isdw*ks*specular*
   * max(0, h.n)^shininess where the gradient is nonzero (ks counts only with shadows
   * on); the light is cvr_light.position.  Off; 0.8, 30, (1, 1, 1). */
  int apply_phong;
  int gradient_type;
  float ks, shininess;
  float specular[3];
  /* Speed and memory only (pixels agree to the last bits): the SAT layout the march
   * reads, 0 corner blocks (64 bytes per cell beside the table, one record per GL_LINEAR
   * tap), 1 the linear (hi, lo) table itself (8 bytes per cell, four 16-byte x-pairs per
   * tap).  0 (53.2 ms against 61.0 ms per 256^3 frame at 1024^2 on MI355X); 1 needs
   * 1/8 of the SAT memory. */
  int sat_layout;
} cvr_ebs_params;

/* RC1PConeLightGroundTruthSteps parameters (crtgtrenderer.cpp:23-56, 106-113).  Each
 * opaque sample casts n rays inside a cone (theta ~ U[0, aperture], phi ~ U[0, 360)
 * drawn once per parameter change from a default-seeded std::default_random_engine,
 * :131-185) and marches them with the trapezoid rule from `initial_gap` in steps of
 * `ray_step` (gt_ray_marching.comp:109-252). */
typedef struct cvr_gt_params {
  float step_size;        /* <= 0: reference default 0.5/sqrt(3)*|voxel scale| */
  float et_threshold;     /* 0.99 (gt_ray_marching.comp:450) */
  int apply_occlusion;    /* m_apply_occlusion (reference default off; here 1) */
  int apply_shadow;       /* m_apply_shadows (reference default off; here 1) */
  int occ_rays;           /* OccNumberOfSampledRays, 1 */
  float occ_aperture_deg; /* OccConeApertureAngle, 90 */
  float occ_max_dist;     /* <= 0: 0.5 * volume diagonal */
  int sdw_rays;           /* SdwNumberOfSampledRays, 1 */
  float sdw_aperture_deg; /* SdwConeApertureAngle, 1 */
  float sdw_max_dist;     /* <= 0: 0.75 * volume diagonal */
  int shadow_type;        /* SdwShadowType: 0 point light (position), 1 spot, 2 directional */
  float initial_gap;      /* LightRayInitialGap, 1.0 */
  float ray_step;         /* LightRayStepSize, 0.5 */
  float ka, kd;           /* 0.5, 0.5 */
} cvr_gt_params;

/* One cone of RC1PConeTracingDirOcclusionShading: the ConeGaussianSampler settings
 * (conegaussiansampler.cpp:30-47, 52-172; out-of-range values are clamped as the
 * reference's setters clamp them). */
typedef struct cvr_dos_cone {
  float half_deg;       /* cone half angle, [0.5, 89.5] */
  int max_packing;      /* 0: one Gaussian per section, 1: up to 3, 2: up to 7 */
  float ui_weight;      /* UIWeight multiplying the optical depth */
  float initial_step;   /* distance of the first section, 3.0 */
  float covered_dist;   /* <= 0: the reference's fraction of the diagonal; at least 10 */
  float d_sigma;        /* integration half-step multiplier, [1, 3], 1.25 */
  float r_sigma;        /* Gaussian sigma limit multiplier, [0.5, 3], 2.0 */
} cvr_dos_cone;

/* RC1PConeTracingDirOcclusionShading parameters (dosrcrenderer.cpp:29-70, 101-132). */
typedef struct cvr_dos_params {
  float step_size;       /* <= 0: reference default */
  float et_threshold;    /* 0.99 (ray_bbox_marching.comp:725) */
  int apply_occlusion;   /* 1 */
  int apply_shadow;      /* 0 */
  int shadow_type;       /* type_of_shadow: 0 point light, 1 spot light, 2 directional */
  float spot_max_deg;    /* spot light half angle (lightsourcelist.cpp:32: 4) */
  int n_levels;          /* pyramid levels kept; <= 0: the full chain down to 1^3 */
  int base_res[3];       /* pyramid level-0 resolution, 128^3 (extcoefvolumegenerator.cpp:10-15); 0: the volume's */
  float sigma0;          /* level-0 Gaussian sigma, 1.0 */
  cvr_dos_cone occ;      /* 20 deg, packing 1 (3 cones), weight 0.35, covered 0.5 * diagonal */
  cvr_dos_cone sdw;      /* 0.5 deg, packing 0, weight 1.0, covered 0.75 * diagonal */
  float ka, kd;          /* 0.5, 0.5 */
  int density_precision; /* pyramid storage read by the cones: 0 fp32, 1 fp16, 2 bf16 (0) */
  /* Object-space light cache as for CVR_EBS (dosrcrenderer.cpp:64, 134-143, 555-651,
   * rc1pdosct/lightcachecomputation.comp:523-546): the occlusion cone of a cell turns
   * towards the eye in a frame built from the camera's up vector.  Off, 32^3. */
  int light_cache;
  int light_cache_res[3];
  /* ApplyPhongShading as for CVR_EBS (ray_bbox_marching.comp:627-650). */
  int apply_phong;
  int gradient_type;
  float ks, shininess;
  float specular[3];
} cvr_dos_params;

/* Cone-traced lighting over the Gaussian extinction pyramid.  The pyramid follows
 * ExtinctionCoefficientVolume (extcoefvolumegenerator.cpp:10-15, 230-408: 7^3
 * Gaussian taps, sigma0 * 2^i per level, extinction = -log(1 - opacity)).  Cone k
 * is marched over n_steps geometric distances t0 * q^j (q = (max/t0)^(1/n_steps));
 * each step reads the level whose sigma is r/r_sigma for the cone radius r there. */
typedef struct cvr_cone_params {
  float step_size;      /* <= 0: reference default */
  float et_threshold;   /* 0.99 */
  int apply_occlusion;  /* 1 */
  int apply_shadow;     /* 1 */
  int n_levels;         /* pyramid levels kept; <= 0: the full chain down to 1^3 */
  int base_res[3];      /* pyramid level-0 resolution; 0: the volume's */
  float sigma0;         /* level-0 Gaussian sigma, world units (1.0) */
  int n_steps;          /* steps per cone (64, at most 256) */
  int n_occ_cones;      /* occlusion cones: 1 axial + n-1 on a ring (8) */
  float occ_half_deg;   /* occlusion cone half angle (30) */
  float ring_deg;       /* angle of the ring cones' axes to the occlusion axis (60) */
  float sdw_half_deg;   /* shadow cone half angle (10) */
  float t0;             /* first distance, world units (1.0) */
  float occ_max_dist;   /* <= 0: 0.5 * diagonal */
  float sdw_max_dist;   /* <= 0: 0.75 * diagonal */
  float r_sigma;        /* cone radius / Gaussian sigma at the level read (2.0) */
  float ka, kd;         /* 0.5, 0.5 */
  int density_precision; /* pyramid storage read by the cones: 0 fp32, 1 fp16, 2 bf16
                            (0; 1 in the surrogate's cvr_nn_params.cone) */
} cvr_cone_params;

/* Surrogate: the cone parameters it replaces (and falls back to) plus training. */
typedef struct cvr_nn_params {
  cvr_cone_params cone;
  int n_train;          /* exact-tracer samples generated for training (2^22) */
  int n_val;            /* held-out samples judging the network (65536) */
  int epochs;           /* 4 */
  int batch;            /* 16384 */
  float lr;             /* peak Adam learning rate, cosine decay (2e-3) */
  float rmse_gate;      /* validation RMSE above which the exact tracer renders (0.05) */
  float ood_margin;     /* inputs this far outside the training range use the exact tracer (0.02) */
  int force_exact;      /* nonzero: every sample uses the exact tracer */
  uint64_t seed;        /* training-set and initialisation seed */
  /* Work shape of the renderer (pixels do not depend on it): wavefront tile width in
   * pixels (4, 8, 16 or 32; the height is 64 / width) and the fraction of a wave's
   * marching lanes that must wait on a shaded sample before the wave shades them
   * together (0: as soon as one waits). */
  int tile_w;           /* 4 */
  float shade_batch;    /* 0.75 */
  int span;             /* wave tiles per wavefront, handed to lanes as their rays end: 1, 2, 4 or 8 (1) */
} cvr_nn_params;

typedef struct cvr_nn_report {
  double train_loss;     /* loss of the last minibatch */
  double val_rmse_occ;   /* validation RMSE of the network vs the exact tracer */
  double val_rmse_sdw;
  double base_rmse_occ;  /* the same for the network-free stencil product */
  double base_rmse_sdw;
  int steps;             /* optimiser steps taken */
  int exact_fallback;    /* 1 when the gate sends every sample to the exact tracer */
} cvr_nn_report;

/* Library / handle lifetime. */
const char* cvr_last_error(void);
const char* cvr_technique_name(int technique);
int cvr_create(int technique, int device, cvr_handle* out);
void cvr_destroy(cvr_handle h);

/* Volume (structuredgridvolume.cpp:121-151): nx*ny*nz voxels, x fastest, uint8 or
 * uint16, voxel scale (sx, sy, sz).  `voxels` is a host pointer unless
 * `is_device` is nonzero.  Replaces DataManager::GenerateStructuredVolumeTexture
 * (datamanager.cpp:311-330) + vis::GenerateRTexture (utils.cpp:20-56). */
int cvr_set_volume(cvr_handle h, int nx, int ny, int nz, int dtype, float sx, float sy, float sz,
                   const void* voxels, int is_device);

/* Transfer function 1D (transferfunction1d.cpp:319-358, reader.cpp:744-807):
 * `rgb_iso` holds n_rgb rows (r, g, b, isovalue), `alpha_iso` n_alpha rows
 * (alpha-or-extinction, isovalue); the table has max_density+1 entries. */
int cvr_set_transfer_function(cvr_handle h, int max_density, int extinction_input, int n_rgb,
                              const double* rgb_iso, int n_alpha, const double* alpha_iso);

/* Technique parameters: a POD of the technique's cvr_*_params type, `bytes` must
 * equal its size.  cvr_default_params fills the reference defaults. */
int cvr_default_params(int technique, void* params, size_t bytes);
int cvr_set_params(cvr_handle h, const void* params, size_t bytes);
int cvr_get_params(cvr_handle h, void* params, size_t bytes);

/* Per-volume/TF preprocessing, the work the reference does in Init (e.g. the EBS
 * SAT, ebsrenderer.cpp:84-123).  Called implicitly by cvr_render when needed. */
int cvr_prepare(cvr_handle h, void* hip_stream);

/* Render one frame (the reference's Update + Redraw, rc1prenderer.cpp:72-151).
 * `out_rgba` receives width*height*4 floats, premultiplied, row 0 at the bottom;
 * pixels whose ray misses the volume are (0,0,0,0).  `out_is_device` selects a
 * device or host destination; `hip_stream` may be NULL (the null stream). */
int cvr_render(cvr_handle h, const cvr_camera* cam, const cvr_light* light, int width, int height,
               float* out_rgba, int out_is_device, void* hip_stream);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc): cvr_render of
 * a CVR_CONE_NN handle into device memory that calls no allocating routine.  The handle
 * must be prepared first (cvr_prepare: pyramid, cone tables and trained network in
 * place; otherwise CVR_ERR_STATE).  The two-pass render's tile marks live in a fixed
 * library array of 65536 words per GPU (256 KiB, held until the process ends), so a
 * frame may have at most 65536 workgroup tiles (4096 x 4096 pixels at the default 4 x 16
 * wave tiles; larger frames give CVR_ERR_INVALID).  Every call on one GPU shares that
 * array: calls that overlap in time on one GPU race.  The pixels equal cvr_render's;
 * stats are not collected. */
int cvr_render_noalloc(cvr_handle h, const cvr_camera* cam, const cvr_light* light, int width, int height,
                       float* out_device, void* hip_stream);

/* The extinction SAT built on the GPU for CVR_EBS ((nx+2)(ny+2)(nz+2) cells,
 * ebsrenderer.cpp:624-723 + summedareatable.h:218-278), copied to the host as
 * double (hi + lo of the two-float storage).  Builds it if needed. */
int cvr_get_sat(cvr_handle h, double* out_host, size_t count);

/* The volume gradient of `type` (built if needed), 3 floats per voxel, x fastest
 * (count = floats available).  Replaces reading back the gradient texture. */
int cvr_get_gradient(cvr_handle h, int type, float* out_host, size_t count);

/* The light cache of the last CVR_EBS / CVR_DOS render with light_cache on: res[3]
 * receives its size, out_host 2 floats (occlusion, shadow) per cell, x fastest
 * (count = floats available).  Replaces reading back the RG16F cache texture. */
int cvr_get_light_cache(cvr_handle h, float* out_host, size_t count, int* res);

/* Multi-scale output (RenderFrameToScreen, libs/vis_utils/renderoutputframe.cpp:96-535):
 * a frame rendered at a multiple (or fraction) of the screen size resampled to it.
 * Kernels follow vis_utils/filters/utils.hpp:8-15 (IMAGE_FILTER_KERNEL, default hat). */
enum {
  CVR_FILTER_BOX = 0,
  CVR_FILTER_HAT = 1,
  CVR_FILTER_CATMULL_ROM = 2,
  CVR_FILTER_MITCHELL_NETRAVALI = 3,
  CVR_FILTER_CARDINAL_BSPLINE3 = 4,
  CVR_FILTER_CARDINAL_OMOMS3 = 5
};
enum {
  CVR_SCALE_MULTISAMPLE = 0, /* DrawMultiSampleHigherResolutionMode: GL_LINEAR at screen pixel centres */
  CVR_SCALE_DOWNSCALE = 1,   /* DrawHigherResolutionWithDownScale: kernel filter (+ digital filter) */
  CVR_SCALE_UPSCALE = 2      /* DrawLowerResolutionWithUpScale: (digital filter +) kernel interpolation */
};
/* The frame size for screen (w, h) and multipliers (mw, mh): m > 0 multiplies, m < 0
 * divides (UpdateScreenResolutionMultiScaling, renderoutputframe.cpp:96-112). */
int cvr_multiscale_size(int screen_w, int screen_h, int mw, int mh, int* frame_w, int* frame_h);
/* Resamples the RGBA float frame src (sw x sh, device) into dst (dw x dh, device) on
 * `device` / hip_stream.  UPSCALE with a cardinal kernel filters a copy of src in
 * `work` (sw*sh*4 floats, device; work == src filters src in place as the reference
 * does); otherwise work may be NULL. */
int cvr_resample_frame(int device, const float* src, int sw, int sh, float* dst, int dw, int dh, int mode, int kernel,
                       float* work, void* hip_stream);

/* Multi-scale rendering of a technique (BaseVolumeRenderer::MULTISCALING and the
 * MultiSampleRedraw / DownScalingRedraw / UpScalingRedraw it selects,
 * cppvolrend/volrenderbase.h:28-33, 50-52, dispatched by RenderingManager::Display,
 * renderingmanager.cpp:204, 1201-1229).  The technique renders a frame of
 * cvr_multiscale_size(screen, +mult) for CVR_MULTIPLE_RAYS_PER_PIXEL and
 * CVR_DOWN_SCALING_RENDER, of cvr_multiscale_size(screen, -mult) for
 * CVR_UP_SCALING_RENDER (Reshape, volrenderbase.cpp:42-66; the reference's multipliers
 * are 2 x 2, defines.h:16-17), which is resampled to the screen size with `kernel`
 * (cvr_resample_frame's CVR_SCALE_MULTISAMPLE / DOWNSCALE / UPSCALE) into out_rgba
 * (screen_w x screen_h RGBA floats, host or device).  CVR_SINGLE_RAY_PER_PIXEL is
 * cvr_render at the screen size.  A camera aspect <= 0 means the screen's w/h.  The
 * handle keeps the multi-scale frame in device memory between calls. */
enum {
  CVR_SINGLE_RAY_PER_PIXEL = 0,
  CVR_MULTIPLE_RAYS_PER_PIXEL = 1,
  CVR_DOWN_SCALING_RENDER = 2,
  CVR_UP_SCALING_RENDER = 3
};
int cvr_render_multiscale(cvr_handle h, const cvr_camera* cam, const cvr_light* light, int screen_w, int screen_h,
                          int mode, int mult_w, int mult_h, int kernel, float* out_rgba, int out_is_device,
                          void* hip_stream);

/* Measurement helper, no reference counterpart: the rate (GB/s of 32-byte reads) at which
 * lanes reading independent 32-byte records at pseudo-random indices get them back from a
 * working set of `working_set_bytes` (16 KiB: L1, 2 MiB: L2, 64 MiB: MALL, 4 GiB: HBM),
 * best of `reps` launches after a warm-up.  The ceiling of the renderers' trilinear
 * corner-block reads from each level, used for their roofline figures. */
int cvr_probe_gather(int device, size_t working_set_bytes, int reps, double* gbps, double* reads_per_launch,
                     void* hip_stream);

/* Counters of the last render (samples marched, shaded samples). */
typedef struct cvr_stats {
  uint64_t samples;          /* volume samples taken along camera rays */
  uint64_t shaded_samples;   /* samples with nonzero extinction (lighting evaluated) */
  uint64_t fallback_samples; /* surrogate: samples the exact tracer shaded */
  double last_ms;            /* device time of the last render's main kernel */
  uint64_t box_sums;         /* CVR_EBS: SAT box sums evaluated (8 trilinear SAT reads each) */
  /* CVR_EBS with stats on: device time of the last SAT build, its prefix passes
   * and the corner-block copy the render kernels read (events around the kernels). */
  double sat_scan_ms;
  double sat_blocks_ms;
} cvr_stats;
int cvr_get_stats(cvr_handle h, cvr_stats* out);
int cvr_set_collect_stats(cvr_handle h, int enable);

/* The extinction pyramid of CVR_DOS / CVR_CONE_TRACE / CVR_CONE_NN (built if needed): level
 * count and dims (3 ints per level, x first), and one level's extinction (fp32,
 * x fastest).  Replaces reading back ExtinctionCoefficientVolume's texture. */
int cvr_pyramid_info(cvr_handle h, int* levels, int* dims, int max_levels);
int cvr_get_pyramid_level(cvr_handle h, int level, float* out_host, size_t count);

/* CVR_DOS: the section table of the occlusion (which = 0) or shadow (which = 1) cone
 * as ConeGaussianSampler::ComputeConeIntegrationSteps builds it in FP64
 * (conegaussiansampler.cpp:212-414) and GetConeSectionsInfoTex lays it out (:179-205):
 * rows (interval to the next section, mip level, d_integral, amplitude) as float.
 * `n_sections` receives the row count (at most max_rows are copied), `counts` the
 * sections with 1, 3 and 7 Gaussians, `axes` the 3 + 7 sub-cone axes (x, y, z) and
 * `adj7` the 7-cone adjacent weight (each output may be NULL). */
int cvr_dos_sections(cvr_handle h, int which, float* rows, int max_rows, int* n_sections, int* counts, float* axes,
                     float* adj7);

/* CVR_CONE_GT: the sampled ray directions of the occlusion (which = 0) or shadow
 * (which = 1) cone, 3 floats each in the cone's (right, up, axis) frame
 * (crtgtrenderer.cpp:131-185).  `n` receives the count. */
int cvr_gt_rays(cvr_handle h, int which, float* out, int max_rays, int* n);

/* Exact cone-traced occlusion and shadow at n points (host arrays): positions in
 * texture space [0, G], camera-ray directions, light directions (3 floats each);
 * outputs occ[n], sdw[n] and the surrogate's 16 inputs feat[16 n] (each may be NULL). */
int cvr_cone_eval(cvr_handle h, int n, const float* pos, const float* dirs, const float* lights, float* occ,
                  float* sdw, float* feat);

/* Surrogate network (CVR_CONE_NN).  cvr_nn_train generates the training set with the
 * exact tracer and trains on the GPU (cvr_prepare does it when needed).  Weights are
 * fp32, cvr_nn_param_count() of them, flattened W1[64][16] b1 W2[64][64] b2
 * W3[64][64] b3 W4[16][64] b4[16]; setting them marks the network trained.
 * cvr_nn_eval runs X[n][16] -> Y[n][2] on the GPU through the training kernels'
 * path (path 0) or the renderer's register-resident path (path 1).
 * cvr_nn_gradient returns dL/dparams of L = sum (y - t)^2 / (2 n) at the current
 * weights, computed by the training kernel.  No reference counterpart. */
int cvr_nn_param_count(void);
int cvr_nn_train(cvr_handle h, void* hip_stream);
int cvr_nn_get_weights(cvr_handle h, float* out, size_t count);
int cvr_nn_set_weights(cvr_handle h, const float* in, size_t count);
int cvr_nn_eval(cvr_handle h, int n, const float* X, float* Y, int path);
int cvr_nn_gradient(cvr_handle h, int n, const float* X, const float* T, float* grad, float* loss);
int cvr_nn_get_report(cvr_handle h, cvr_nn_report* out);

/* Checkpoint of a trained surrogate: a versioned binary holding the fp32 weights, the
 * Adam moments, the training-set input range and the report, so a network is restored
 * exactly (and training could resume).  cvr_nn_checkpoint_size gives the byte count. */
size_t cvr_nn_checkpoint_size(void);
int cvr_nn_save(cvr_handle h, void* out, size_t bytes);
int cvr_nn_load(cvr_handle h, const void* in, size_t bytes);

/* ---------------------------------------------------------------------------
 * Batched rendering.  No reference counterpart: the reference renders one volume
 * from one view per dispatch (rc1prenderer.cpp:140-151, renderingmanager.cpp:174-319).
 * A batch references prepared-on-demand CVR_CONE_NN handles (each owns its volume,
 * transfer function, pyramid and trained surrogate; the batch does not own them and
 * they must outlive it) and a list of jobs (volume, camera, light) at one image size.
 * cvr_batch_render renders every job in one persistent launch: workgroups pull
 * (job, tile) items from device work queues (one per XCD, with stealing, or one
 * global queue); the queue reset and the launch are captured once into a HIP graph
 * and replayed while jobs, tuning and output stay the same.  Images land in `out`
 * (device memory, n_jobs * height * width * 4 floats, job-major, each image as
 * cvr_render writes it).  Each image equals cvr_render of the same handle, camera and
 * light up to fp32 rounding (the same device code compiled into another kernel), and
 * does not depend on the work shape.
 */
typedef struct cvr_batch* cvr_batch_handle;

typedef struct cvr_job {
  int volume;        /* index returned by cvr_batch_add_volume */
  cvr_camera camera;
  cvr_light light;
} cvr_job;

/* Work shape of the persistent launch and step schedule.  The pixels do not depend on
 * the work shape (tile_w .. span); step_scale changes the march step and with it the
 * pixels, within the image error the autotuner accepts for it. */
typedef struct cvr_batch_tuning {
  int tile_w;            /* wavefront tile width: 4, 8, 16 or 32 pixels (height 64 / width) */
  float shade_batch;     /* fraction of marching lanes that wait before a shading pass, [0, 1] */
  int queues;            /* 1: one global queue; 8: one queue per XCD with stealing */
  int workgroups_per_cu; /* persistent workgroups per compute unit, 1..4 */
  int span;              /* wave tiles per wavefront, handed to lanes as their rays end: 1, 2, 4 or 8 */
  double ms;             /* autotune: the measured time of the chosen shape */
  float step_scale;      /* multiplier of every member's march step (its cone.step_size, or the
                            default half-voxel step), [0.25, 4]; 1 = each member's own step.  The
                            reference sweeps this step by hand (rc1prenderer.cpp:225-229). */
  float max_step_rmse;   /* autotune input: the largest image RMSE (RGB of the pixels either image
                            covers) a step_scale above 1 may add against step_scale 1 on the tuned
                            jobs; 0 (default) leaves step_scale unchanged */
  double step_rmse;      /* autotune output: that RMSE for the chosen step_scale (0 when not tuned) */
} cvr_batch_tuning;

int cvr_batch_create(int device, int width, int height, cvr_batch_handle* out);
void cvr_batch_destroy(cvr_batch_handle b);
int cvr_batch_add_volume(cvr_batch_handle b, cvr_handle h, int* index);
int cvr_batch_set_jobs(cvr_batch_handle b, int n, const cvr_job* jobs);
int cvr_batch_set_tuning(cvr_batch_handle b, const cvr_batch_tuning* t);
int cvr_batch_get_tuning(cvr_batch_handle b, cvr_batch_tuning* t);
/* Times candidate shapes on every `stride`-th job (stride >= 1) rendered into `out`,
 * keeps the fastest and reports it.  With max_step_rmse > 0 in the current tuning it
 * first searches the step schedule: step_scale 1, 1.25, 1.5, 2, 2.5, 3 in turn, each
 * rendered and compared on the device with the step_scale 1 images, up to the first that
 * adds more than max_step_rmse; the fastest admitted step is kept, then the shape is
 * tuned at that step.  Allocates one temporary image buffer for the tuned jobs. */
int cvr_batch_autotune(cvr_batch_handle b, float* out, int stride, void* hip_stream, cvr_batch_tuning* best);
int cvr_batch_render(cvr_batch_handle b, float* out, void* hip_stream);
/* Counters of the last batch render (waits for the device). */
int cvr_batch_get_stats(cvr_batch_handle b, cvr_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* CVR_H */
