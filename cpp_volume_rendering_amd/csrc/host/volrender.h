// Headless plugin surface mirroring the reference's BaseVolumeRenderer
// (cppvolrend/volrenderbase.h:25-98) and RenderingManager
// (cppvolrend/renderingmanager.h:29-167).  Each technique class is a thin shim over
// one cvr_handle of the C ABI; the only interface change forced by dropping OpenGL is
// GetScreenTextureID() -> GetFrame() (a host RGBA float buffer, row 0 at the bottom).
#pragma once
#include <string>
#include <vector>

#include "vis.h"

class BaseVolumeRenderer {
 public:
  // volrenderbase.h:28-33
  enum MULTISCALING {
    SINGLE_RAY_PER_PIXEL = CVR_SINGLE_RAY_PER_PIXEL,
    MULTIPLE_RAYS_PER_PIXEL = CVR_MULTIPLE_RAYS_PER_PIXEL,
    DOWN_SCALING_RENDER = CVR_DOWN_SCALING_RENDER,
    UP_SCALING_RENDER = CVR_UP_SCALING_RENDER,
  };
  // MULTISAMPLE_NUMBEROFSAMPLES_W / _H (defines.h:16-17)
  static constexpr int kMultiSampleW = 2, kMultiSampleH = 2;

  BaseVolumeRenderer();
  virtual ~BaseVolumeRenderer();

  // volrenderbase.cpp:20-24: non-owning external resources
  void SetExternalResources(vis::DataManager* data_mgr, vis::RenderingParameters* rdr_prm);

  virtual const char* GetName() = 0;
  virtual const char* GetAbbreviationName() = 0;

  virtual void Clean();
  virtual void ReloadShaders() {}
  virtual bool Init(int shader_width, int shader_height) = 0;
  virtual bool Update(vis::Camera* camera) = 0;
  virtual void Redraw() = 0;
  // volrenderbase.h:50-52: the frame rendered at the mode's size and resampled to the
  // screen (the base versions do nothing, as the reference's)
  virtual void MultiSampleRedraw() {}
  virtual void DownScalingRedraw() {}
  virtual void UpScalingRedraw() {}
  // volrenderbase.cpp:42-66: the render size follows the multi-scaling mode
  virtual void Reshape(int w, int h);
  virtual vis::GRID_VOLUME_DATA_TYPE GetDataTypeSupport() { return vis::GRID_VOLUME_DATA_TYPE::STRUCTURED; }

  // volrenderbase.cpp:77-84
  void PrepareRender(vis::Camera* camera);
  virtual void SetOutdated() { vr_outdated = true; }
  bool IsOutdated() const { return vr_outdated; }
  bool IsBuilt() const { return vr_built; }

  // volrenderbase.cpp:103-121
  bool IsPixelMultiScalingSupported() const { return vr_pixel_multiscaling_support; }
  int GetCurrentMultiScalingMode() const { return IsPixelMultiScalingSupported() ? vr_pixel_multiscaling_mode : 0; }
  void SetCurrentMultiScalingMode(int f);
  // RenderFrameToScreen::Get/SetImageKernelFilter (renderoutputframe.h:53-57): a CVR_FILTER_* kernel
  unsigned int GetImageKernelFilter() const { return m_kernel_filter; }
  void SetImageKernelFilter(unsigned int k);
  // The size the technique renders at (the screen size, or its multiple / fraction).
  int GetRenderWidth() const { return m_render_w; }
  int GetRenderHeight() const { return m_render_h; }

  // The rendered frame: width * height premultiplied RGBA floats, row 0 at the bottom.
  const std::vector<float>& GetFrame() const { return m_frame; }
  int GetFrameWidth() const { return m_width; }
  int GetFrameHeight() const { return m_height; }
  // Status of the last failing call of the C ABI (empty when none).
  const std::string& GetLastError() const { return m_error; }

 protected:
  void SetBuilt(bool b) { vr_built = b; }

  bool vr_built = false, vr_outdated = true;
  bool vr_pixel_multiscaling_support = false;
  int vr_pixel_multiscaling_mode = 0;
  unsigned int m_kernel_filter = CVR_FILTER_HAT;  // IMAGE_FILTER_KERNEL default (filters/utils.hpp:8-15)
  vis::DataManager* m_ext_data_manager = nullptr;
  vis::RenderingParameters* m_ext_rendering_parameters = nullptr;
  int m_width = 0, m_height = 0;      // screen
  int m_render_w = 0, m_render_h = 0;  // frame the technique renders
  std::vector<float> m_frame;
  std::string m_error;
};

// One technique of the C ABI behind the plugin surface.
class GpuVolumeRenderer : public BaseVolumeRenderer {
 public:
  explicit GpuVolumeRenderer(int technique, int device = 0);
  ~GpuVolumeRenderer() override;
  void Clean() override;
  bool Init(int shader_width, int shader_height) override;
  bool Update(vis::Camera* camera) override;
  void Redraw() override;
  void MultiSampleRedraw() override;
  void DownScalingRedraw() override;
  void UpScalingRedraw() override;
  // Technique parameters (the cvr_*_params POD of this technique).
  bool SetParams(const void* params, size_t bytes);
  bool GetParams(void* params, size_t bytes);
  cvr_handle Handle() const { return m_handle; }

 protected:
  bool Check(int rc);
  void MultiScaleRedraw(int mode);
  int m_technique, m_device;
  cvr_handle m_handle = nullptr;
  cvr_camera m_cam{};
  cvr_light m_light{};
};

// RayCasting1Pass (rc1prenderer.h:32-78)
class RayCasting1Pass : public GpuVolumeRenderer {
 public:
  RayCasting1Pass() : GpuVolumeRenderer(CVR_RC1P) {}
  const char* GetName() override { return "1-Pass - Ray Casting"; }
  const char* GetAbbreviationName() override { return "s_1rc"; }
};

// RC1PExtinctionBasedShading (ebsrenderer.h:45-119)
class RC1PExtinctionBasedShading : public GpuVolumeRenderer {
 public:
  RC1PExtinctionBasedShading() : GpuVolumeRenderer(CVR_EBS) {}
  const char* GetName() override { return "1-Pass - Ray Casting - Extinction-based"; }
  const char* GetAbbreviationName() override { return "s_1rc_eb"; }
};

// RC1PConeLightGroundTruthSteps (crtgtrenderer.h:50-120): ray-sampled cone ground truth
class RC1PConeLightGroundTruthSteps : public GpuVolumeRenderer {
 public:
  RC1PConeLightGroundTruthSteps() : GpuVolumeRenderer(CVR_CONE_GT) {}
  const char* GetName() override { return "1-Pass - Ray Casting - Cone Ground Truth (Steps)"; }
  const char* GetAbbreviationName() override { return "s_1rc_gt_c"; }
};

// RC1PConeTracingDirOcclusionShading (dosrcrenderer.h:38-121): cone section tables
// integrated over the Gaussian extinction pyramid
class RC1PConeTracingDirOcclusionShading : public GpuVolumeRenderer {
 public:
  RC1PConeTracingDirOcclusionShading() : GpuVolumeRenderer(CVR_DOS) {}
  const char* GetName() override { return "1-Pass - Ray Casting - Dir. Occlusion Shading"; }
  const char* GetAbbreviationName() override { return "s_1rc_dos"; }
};

// Cone-traced directional occlusion + shadows over the Gaussian extinction pyramid
// (the pyramid and shading of RC1PConeTracingDirOcclusionShading, dosrcrenderer.h:38-121)
class RC1PConeTracedOcclusion : public GpuVolumeRenderer {
 public:
  RC1PConeTracedOcclusion() : GpuVolumeRenderer(CVR_CONE_TRACE) {}
  const char* GetName() override { return "1-Pass - Ray Casting - Cone-traced Directional Occlusion"; }
  const char* GetAbbreviationName() override { return "s_1rc_ct"; }
};

// The same renderer with the learned surrogate (no reference counterpart).
class RC1PLearnedConeOcclusion : public GpuVolumeRenderer {
 public:
  RC1PLearnedConeOcclusion() : GpuVolumeRenderer(CVR_CONE_NN) {}
  const char* GetName() override { return "1-Pass - Ray Casting - Learned Cone Occlusion"; }
  const char* GetAbbreviationName() override { return "s_1rc_nn"; }
};

// Headless RenderingManager: owns the renderers (renderingmanager.cpp:119-123,
// 372-379), the data manager, camera and rendering parameters; Display() drives one
// frame (:174-319); EvaluateFrames() is the eval.csv harness (:261-317, 807-868).
class RenderingManager {
 public:
  static RenderingManager* Instance();
  static void DestroyInstance();

  void AddVolumeRenderer(BaseVolumeRenderer* vr);  // takes ownership
  int GetNumberOfVolumeRenderers() const { return int(m_renderers.size()); }
  BaseVolumeRenderer* GetVolumeRenderer(int i) const { return m_renderers[i]; }
  // renderingmanager.cpp:521-540; false if the technique cannot handle the data
  bool SetCurrentVolumeRenderer(int id);
  BaseVolumeRenderer* GetCurrentVolumeRenderer() const { return m_current; }

  // renderingmanager.cpp:126-172 with the volume and transfer function given directly.
  bool InitData(vis::StructuredGridVolume* vol, vis::TransferFunction1D* tf, int width, int height);
  // Light axes from the camera basis (renderingmanager.cpp:168, 421-426), or a direction.
  void UpdateLightSourceCameraVectors();
  void SetLightDirection(const std::array<float, 3>& towards_light);

  bool Display();
  bool SaveFramePFM(const std::string& path) const;  // float RGBA -> PFM (RGB)
  bool SaveFramePPM(const std::string& path) const;  // 8-bit RGB over a white background
  // Renders `frames` frames of the current renderer; appends "name,ms,fps" to `csv`.
  bool EvaluateFrames(int frames, const std::string& csv, double* ms_per_frame = nullptr);

  vis::Camera& GetCamera() { return m_camera; }
  vis::RenderingParameters& GetRenderingParameters() { return m_params; }
  vis::DataManager& GetDataManager() { return m_data; }
  void CloseFunc();

 private:
  RenderingManager() = default;
  ~RenderingManager();
  std::vector<BaseVolumeRenderer*> m_renderers;
  BaseVolumeRenderer* m_current = nullptr;
  vis::DataManager m_data;
  vis::RenderingParameters m_params;
  vis::Camera m_camera;
  static RenderingManager* s_instance;
};
