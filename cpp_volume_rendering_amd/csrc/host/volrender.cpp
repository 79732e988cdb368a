// Headless plugin surface (declarations: volrender.h).
#include "volrender.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>

// ------------------------------------------------------------------ BaseVolumeRenderer
BaseVolumeRenderer::BaseVolumeRenderer() {}
BaseVolumeRenderer::~BaseVolumeRenderer() {}

void BaseVolumeRenderer::SetExternalResources(vis::DataManager* data_mgr, vis::RenderingParameters* rdr_prm) {
  m_ext_data_manager = data_mgr;
  m_ext_rendering_parameters = rdr_prm;
}

void BaseVolumeRenderer::Clean() {
  vr_built = false;
  vr_outdated = true;
}

void BaseVolumeRenderer::Reshape(int w, int h) {
  int rw = w, rh = h;
  const int mode = GetCurrentMultiScalingMode();
  if (mode == MULTIPLE_RAYS_PER_PIXEL || mode == DOWN_SCALING_RENDER)
    (void)cvr_multiscale_size(w, h, kMultiSampleW, kMultiSampleH, &rw, &rh);
  else if (mode == UP_SCALING_RENDER)
    (void)cvr_multiscale_size(w, h, -kMultiSampleW, -kMultiSampleH, &rw, &rh);
  m_render_w = rw;
  m_render_h = rh;
  if (w != m_width || h != m_height) {
    m_width = w;
    m_height = h;
    m_frame.assign(size_t(w) * h * 4, 0.f);
  }
  SetOutdated();
}

void BaseVolumeRenderer::SetCurrentMultiScalingMode(int f) {
  vr_pixel_multiscaling_mode = f;
  if (m_width > 0) Reshape(m_width, m_height);
}

void BaseVolumeRenderer::SetImageKernelFilter(unsigned int k) {
  m_kernel_filter = k;
  SetOutdated();
}

void BaseVolumeRenderer::PrepareRender(vis::Camera* camera) {
  if (IsOutdated()) {
    Update(camera);
    vr_outdated = false;
  }
}

// ------------------------------------------------------------------ GpuVolumeRenderer
// Every technique renders at any frame size, so all support the multi-scale modes (the
// reference's structured renderers set vr_pixel_multiscaling_support, e.g.
// rc1prenderer.cpp:24-26).
GpuVolumeRenderer::GpuVolumeRenderer(int technique, int device) : m_technique(technique), m_device(device) {
  vr_pixel_multiscaling_support = true;
}

GpuVolumeRenderer::~GpuVolumeRenderer() { Clean(); }

bool GpuVolumeRenderer::Check(int rc) {
  if (rc == CVR_OK) return true;
  m_error = std::string("cvr error ") + std::to_string(rc) + ": " + cvr_last_error();
  return false;
}

void GpuVolumeRenderer::Clean() {
  if (m_handle) cvr_destroy(m_handle);
  m_handle = nullptr;
  BaseVolumeRenderer::Clean();
}

// rc1prenderer.cpp:50-70: fails when there is no volume; uploads volume + TF, sizes the frame.
bool GpuVolumeRenderer::Init(int w, int h) {
  if (IsBuilt()) Clean();
  m_error.clear();
  if (!m_ext_data_manager || !m_ext_data_manager->GetCurrentStructuredVolume() ||
      !m_ext_data_manager->GetCurrentTransferFunction()) {
    m_error = "no volume / transfer function";
    return false;
  }
  vis::StructuredGridVolume* v = m_ext_data_manager->GetCurrentStructuredVolume();
  vis::TransferFunction1D* tf = m_ext_data_manager->GetCurrentTransferFunction();
  if (!Check(cvr_create(m_technique, m_device, &m_handle))) return false;
  if (!Check(cvr_set_volume(m_handle, v->GetWidth(), v->GetHeight(), v->GetDepth(),
                            v->GetBytesPerValue() == 1 ? CVR_U8 : CVR_U16, float(v->GetScaleX()),
                            float(v->GetScaleY()), float(v->GetScaleZ()), v->GetArrayData(), 0)))
    return false;
  const std::vector<double>& rgb = tf->RGBIso();
  const std::vector<double>& a = tf->AlphaIso();
  if (!Check(cvr_set_transfer_function(m_handle, tf->GetMaxDensity(), tf->IsExtinctionCoefficientInput() ? 1 : 0,
                                       int(rgb.size() / 4), rgb.data(), int(a.size() / 2), a.data())))
    return false;
  Reshape(w, h);
  SetBuilt(true);
  SetOutdated();
  return true;
}

bool GpuVolumeRenderer::SetParams(const void* params, size_t bytes) {
  if (!m_handle) { m_error = "not initialised"; return false; }
  SetOutdated();
  return Check(cvr_set_params(m_handle, params, bytes));
}

bool GpuVolumeRenderer::GetParams(void* params, size_t bytes) {
  if (!m_handle) { m_error = "not initialised"; return false; }
  return Check(cvr_get_params(m_handle, params, bytes));
}

// Update (rc1prenderer.cpp:72-138): camera and light uniforms; preprocessing on demand.
bool GpuVolumeRenderer::Update(vis::Camera* camera) {
  if (!m_handle) return false;
  camera->UpdateAspectRatio(float(m_width), float(m_height));
  m_cam = camera->ToC();
  if (m_ext_rendering_parameters) m_light = m_ext_rendering_parameters->ToLight();
  return Check(cvr_prepare(m_handle, nullptr));
}

// Redraw (rc1prenderer.cpp:140-151): clear + dispatch; the frame lands in GetFrame().
void GpuVolumeRenderer::Redraw() {
  if (!m_handle) return;
  Check(cvr_render(m_handle, &m_cam, m_ext_rendering_parameters ? &m_light : nullptr, m_width, m_height,
                   m_frame.data(), 0, nullptr));
}

// MultiSampleRedraw / DownScalingRedraw / UpScalingRedraw (rc1prenderer.cpp:153-190): the
// dispatch at the render size, then RenderFrameToScreen's draw of the mode; the
// screen-sized result lands in GetFrame().
void GpuVolumeRenderer::MultiScaleRedraw(int mode) {
  if (!m_handle) return;
  Check(cvr_render_multiscale(m_handle, &m_cam, m_ext_rendering_parameters ? &m_light : nullptr, m_width, m_height,
                              mode, kMultiSampleW, kMultiSampleH, int(m_kernel_filter), m_frame.data(), 0, nullptr));
}
void GpuVolumeRenderer::MultiSampleRedraw() { MultiScaleRedraw(MULTIPLE_RAYS_PER_PIXEL); }
void GpuVolumeRenderer::DownScalingRedraw() { MultiScaleRedraw(DOWN_SCALING_RENDER); }
void GpuVolumeRenderer::UpScalingRedraw() { MultiScaleRedraw(UP_SCALING_RENDER); }

// ------------------------------------------------------------------ RenderingManager
RenderingManager* RenderingManager::s_instance = nullptr;

RenderingManager* RenderingManager::Instance() {
  if (!s_instance) s_instance = new RenderingManager();
  return s_instance;
}

void RenderingManager::DestroyInstance() {
  delete s_instance;
  s_instance = nullptr;
}

RenderingManager::~RenderingManager() { CloseFunc(); }

void RenderingManager::CloseFunc() {
  for (BaseVolumeRenderer* r : m_renderers) delete r;
  m_renderers.clear();
  m_current = nullptr;
}

void RenderingManager::AddVolumeRenderer(BaseVolumeRenderer* vr) {
  vr->SetExternalResources(&m_data, &m_params);
  m_renderers.push_back(vr);
}

bool RenderingManager::SetCurrentVolumeRenderer(int id) {
  if (id < 0 || id >= int(m_renderers.size())) return false;
  BaseVolumeRenderer* r = m_renderers[id];
  if (r->GetDataTypeSupport() != vis::GRID_VOLUME_DATA_TYPE::STRUCTURED) return false;
  if (m_current && m_current != r) m_current->Clean();
  m_current = r;
  if (m_data.GetCurrentStructuredVolume())
    return r->Init(m_params.GetScreenWidth(), m_params.GetScreenHeight());
  return true;
}

// The camera starts on the orbit the reference's arcball starts from (centre + radius
// along z, camera.cpp:61-65), at twice the volume diagonal.
bool RenderingManager::InitData(vis::StructuredGridVolume* vol, vis::TransferFunction1D* tf, int width, int height) {
  m_data.SetStructuredVolume(vol);
  m_data.SetTransferFunction(tf);
  m_params.SetScreenSize(width, height);
  const float r = float(2.0 * vol->GetDiagonal());
  m_camera.SetData({0.f, 0.f, r}, {0.f, 0.f, 0.f}, {0.f, 1.f, 0.f});
  UpdateLightSourceCameraVectors();
  if (m_current) return m_current->Init(width, height);
  return true;
}

void RenderingManager::UpdateLightSourceCameraVectors() {
  std::array<float, 3> f, u, r;
  m_camera.GetCameraVectors(&f, &u, &r);
  m_params.SetBlinnPhongLightSourceCameraVectors(f, u, r);
  m_params.SetBlinnPhongLightingPosition(m_camera.GetEye());
}

void RenderingManager::SetLightDirection(const std::array<float, 3>& d) {
  vis::Camera lc;
  const float r = m_data.GetCurrentStructuredVolume() ? float(2.0 * m_data.GetCurrentStructuredVolume()->GetDiagonal()) : 1.f;
  const double l = std::sqrt(double(d[0]) * d[0] + double(d[1]) * d[1] + double(d[2]) * d[2]);
  std::array<float, 3> eye = {float(d[0] / l * r), float(d[1] / l * r), float(d[2] / l * r)};
  const bool vertical = std::fabs(d[1] / l) > 0.99;
  lc.SetData(eye, {0.f, 0.f, 0.f}, vertical ? std::array<float, 3>{0.f, 0.f, 1.f} : std::array<float, 3>{0.f, 1.f, 0.f});
  std::array<float, 3> f, u, rr;
  lc.GetCameraVectors(&f, &u, &rr);
  m_params.SetBlinnPhongLightSourceCameraVectors(f, u, rr);
  m_params.SetBlinnPhongLightingPosition(eye);
}

// renderingmanager.cpp:199-208 with f_render (:1201-1229): the redraw of the current
// multi-scaling mode.
bool RenderingManager::Display() {
  if (!m_current || !m_current->IsBuilt()) return false;
  m_current->PrepareRender(&m_camera);
  switch (m_current->GetCurrentMultiScalingMode()) {
    case BaseVolumeRenderer::MULTIPLE_RAYS_PER_PIXEL: m_current->MultiSampleRedraw(); break;
    case BaseVolumeRenderer::DOWN_SCALING_RENDER: m_current->DownScalingRedraw(); break;
    case BaseVolumeRenderer::UP_SCALING_RENDER: m_current->UpScalingRedraw(); break;
    default: m_current->Redraw();
  }
  return m_current->GetLastError().empty();
}

bool RenderingManager::SaveFramePFM(const std::string& path) const {
  if (!m_current) return false;
  const int w = m_current->GetFrameWidth(), h = m_current->GetFrameHeight();
  std::ofstream o(path, std::ios::binary);
  if (!o) return false;
  o << "PF\n" << w << " " << h << "\n-1.0\n";  // little endian, rows bottom to top
  const std::vector<float>& f = m_current->GetFrame();
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) o.write(reinterpret_cast<const char*>(&f[(size_t(y) * w + x) * 4]), 3 * sizeof(float));
  return bool(o);
}

bool RenderingManager::SaveFramePPM(const std::string& path) const {
  if (!m_current) return false;
  const int w = m_current->GetFrameWidth(), h = m_current->GetFrameHeight();
  std::ofstream o(path, std::ios::binary);
  if (!o) return false;
  o << "P6\n" << w << " " << h << "\n255\n";
  const std::vector<float>& f = m_current->GetFrame();
  for (int y = h - 1; y >= 0; --y)  // PPM rows run top to bottom
    for (int x = 0; x < w; ++x) {
      const float* p = &f[(size_t(y) * w + x) * 4];
      for (int c = 0; c < 3; ++c) {
        const float v = p[c] + (1.0f - p[3]);  // premultiplied over white (renderingmanager.cpp:109)
        o.put(char(int(std::lround(std::fmin(std::fmax(v, 0.f), 1.f) * 255.f))));
      }
    }
  return bool(o);
}

bool RenderingManager::EvaluateFrames(int frames, const std::string& csv, double* ms_per_frame) {
  if (!m_current || frames < 1) return false;
  if (!Display()) return false;  // warm-up (and preprocessing)
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < frames; ++i)
    if (!Display()) return false;
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / frames;
  if (ms_per_frame) *ms_per_frame = ms;
  if (!csv.empty()) {
    std::ofstream o(csv, std::ios::app);
    o << m_current->GetAbbreviationName() << "," << ms << "," << 1000.0 / ms << "\n";
  }
  return true;
}
