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
 */
enum cvr_technique { CVR_RC1P = 0, CVR_EBS = 1 };

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
typedef struct cvr_rc1p_params {
  float step_size;    /* <= 0: reference default 0.5/sqrt(3)*|voxel scale| */
  float et_threshold; /* early termination, ray_marching_1p.comp:238 (0.99) */
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
} cvr_ebs_params;

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

/* The extinction SAT built on the GPU for CVR_EBS ((nx+2)(ny+2)(nz+2) cells,
 * ebsrenderer.cpp:624-723 + summedareatable.h:218-278), copied to the host as
 * double (hi + lo of the two-float storage).  Builds it if needed. */
int cvr_get_sat(cvr_handle h, double* out_host, size_t count);

/* Counters of the last render (samples marched, shaded samples). */
typedef struct cvr_stats {
  uint64_t samples;        /* volume samples taken along camera rays */
  uint64_t shaded_samples; /* samples with nonzero extinction (lighting evaluated) */
  double last_ms;          /* device time of the last render's main kernel */
} cvr_stats;
int cvr_get_stats(cvr_handle h, cvr_stats* out);
int cvr_set_collect_stats(cvr_handle h, int enable);

#ifdef __cplusplus
}
#endif
#endif /* CVR_H */
