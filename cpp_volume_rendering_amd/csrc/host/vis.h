// Headless host data model mirroring the reference's libs/volvis_utils and libs/vis_utils:
// structured volume, 1D transfer function and its .tf1d reader, the structured volume
// readers (.raw, .syn, .nrrd/.nhrd, .dat), the dataset list provider, camera,
// rendering parameters and the data manager.  No OpenGL: GPU resources live behind the
// C ABI (include/cvr.h).
#pragma once
#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/cvr.h"

namespace vis {

enum class GRID_VOLUME_DATA_TYPE { STRUCTURED = 0, UNSTRUCTURED = 1 };

// libs/volvis_utils/structuredgridvolume.{h,cpp}: x-fastest voxels, 8 or 16 bit.
class StructuredGridVolume {
 public:
  StructuredGridVolume(std::string name, int w, int h, int d);
  void SetName(const std::string& n) { name_ = n; }
  const std::string& GetName() const { return name_; }
  void SetScale(double sx, double sy, double sz);
  double GetScaleX() const { return scale_[0]; }
  double GetScaleY() const { return scale_[1]; }
  double GetScaleZ() const { return scale_[2]; }
  std::array<double, 3> GetScale() const { return scale_; }
  int GetWidth() const { return w_; }
  int GetHeight() const { return h_; }
  int GetDepth() const { return d_; }
  // takes ownership of the voxel bytes; bytes_per_value 1 or 2
  bool SetArrayData(std::vector<uint8_t> bytes, int bytes_per_value);
  int GetBytesPerValue() const { return bpv_; }
  const void* GetArrayData() const { return data_.empty() ? nullptr : data_.data(); }
  // GetNormalizedSample (structuredgridvolume.cpp:121-151): v/255 or v/65535, 0 outside
  double GetNormalizedSample(int x, int y, int z) const;
  // GetDiagonal (structuredgridvolume.cpp:96-102)
  double GetDiagonal() const;

 private:
  std::string name_;
  int w_, h_, d_, bpv_ = 0;
  std::array<double, 3> scale_{1.0, 1.0, 1.0};
  std::vector<uint8_t> data_;
};

// libs/volvis_utils/transferfunction1d.{h,cpp}: piecewise-linear RGB and alpha (or
// extinction) control points over isovalues [0, max_density].
class TransferFunction1D {
 public:
  explicit TransferFunction1D(int max_value = 255);
  void SetName(const std::string& n) { name_ = n; }
  const std::string& GetName() const { return name_; }
  void SetExtinctionCoefficientInput(bool s) { extinction_coef_type_ = s; built_ = false; }
  bool IsExtinctionCoefficientInput() const { return extinction_coef_type_; }
  void AddRGBControlPoint(double r, double g, double b, int isovalue);
  void AddAlphaControlPoint(double alpha, int isovalue);
  void ClearControlPoints();
  int GetMaxDensity() const { return max_density_; }
  // BuildLinear (transferfunction1d.cpp:319-358); false on repeated/out-of-range isovalues
  bool Build();
  // Get (transferfunction1d.cpp:132-157), GetExtN (:189-197)
  std::array<float, 4> Get(double value, double max_data_value = -1.0);
  float GetExtN(double normalized_value);
  // control points as the C ABI takes them: rows (r,g,b,iso) and (a,iso)
  const std::vector<double>& RGBIso() const { return rgb_iso_; }
  const std::vector<double>& AlphaIso() const { return alpha_iso_; }
  const std::vector<double>& Table() { if (!built_) Build(); return table_; }

 private:
  std::string name_;
  int max_density_;
  bool extinction_coef_type_ = false, built_ = false;
  std::vector<double> rgb_iso_, alpha_iso_, table_;
};

// TransferFunctionReader::readtf1d (reader.cpp:744-807).  Returns nullptr (and an
// error message) instead of exiting on malformed input.
class TransferFunctionReader {
 public:
  static TransferFunction1D* ReadTransferFunction(const std::string& file, std::string* err = nullptr);
  static TransferFunction1D* readtf1d(const std::string& file, std::string* err = nullptr);
};

// VolumeReader (reader.cpp:29-60): dispatch on the extension.  Every reader returns
// nullptr and an error message instead of printing and continuing on bad input.
//  - readraw (reader.cpp:162-225): <name>.<bytes>.<W>x<H>x<D>.raw, voxel scale 1.
//  - readsyn (reader.cpp:283-371): text "W H D", then records "1 x0 y0 z0 x1 y1 z1 v"
//    (fill the half-open box with v) or "<other> x y z v" (one voxel); 8 bit, scale 1.
//  - readnrrd / readnhrd (reader.cpp:373-496): "key: value" header; sizes, type,
//    data file (relative to the header's folder); scale 1.
//  - readdat (reader.cpp:498-628): ObjectFileName, Resolution, SliceThickness (the
//    voxel scale), Format UCHAR / USHORT.
// .pvm / .pvmold (a third-party compressed format) are rejected.
class VolumeReader {
 public:
  static StructuredGridVolume* ReadStructuredVolume(const std::string& file, std::string* err = nullptr);
  static StructuredGridVolume* readraw(const std::string& file, std::string* err = nullptr);
  static StructuredGridVolume* readsyn(const std::string& file, std::string* err = nullptr);
  static StructuredGridVolume* readnrrd(const std::string& file, std::string* err = nullptr);
  static StructuredGridVolume* readnhrd(const std::string& file, std::string* err = nullptr);
  static StructuredGridVolume* readdat(const std::string& file, std::string* err = nullptr);
  // SetArrayDataFromRawFile (reader.cpp:62-98): W*H*D values of 1 or 2 bytes, x fastest
  static bool SetArrayDataFromRawFile(const std::string& file, StructuredGridVolume* sg, int bytes_per_value,
                                      std::string* err = nullptr, size_t offset = 0);
};

// DataProvider (volvis_utils/dataprovider.cpp:27-215): a list of structured datasets,
// either the "<path> <ui name>" list (data/#list_structured_datasets) or a CSV with
// an optional "sep=;" line and the columns Name, Source, File, Size, Spacing, Type
// (data/volume_list.csv).  Paths are relative to the list's folder.
class DataProvider {
 public:
  void ClearStructuredGridFileList();
  // file_path is the folder (with or without a trailing '/'), file_name the list
  bool SetStructuredGridFileList(const std::string& file_path, const std::string& file_name, std::string* err = nullptr);
  int GetNumberOfStructuredGrids() const { return int(names_.size()); }
  int FindStructuredGridName(const std::string& name) const;
  const std::vector<std::string>& GetStructuredGridNameList() const { return names_; }
  StructuredGridVolume* LoadStructuredGrid(unsigned int id, std::string* err = nullptr);
  static StructuredGridVolume* LoadStructuredGridFromRawFile(const std::string& name, const std::string& raw_file_path,
                                                            const std::array<int, 3>& resolution,
                                                            const std::array<double, 3>& scale, int bytes_per_value,
                                                            std::string* err = nullptr);
  static std::string GetExtension(const std::string& filename);
  static int Split(std::vector<std::string>& tags, const std::string& line, const std::string& separator);
  // uchar / uint8 -> 1, ushort / uint16 -> 2, anything else -> 0
  static int GetBytesPerValueByType(const std::string& value_type);

 private:
  std::string dir_, list_;
  bool csv_ = false;
  std::string sep_ = ";";
  std::vector<std::string> tags_, names_, paths_;  // paths_: list format only
  std::vector<std::vector<std::string>> rows_;     // CSV rows
};

// vis_utils/camera.cpp: lookAt camera, fovy 45 degrees, aspect w/h.
class Camera {
 public:
  Camera();
  void SetData(const std::array<float, 3>& eye, const std::array<float, 3>& center, const std::array<float, 3>& up);
  void UpdateAspectRatio(float w, float h) { aspect_ = w / h; }
  float GetAspectRatio() const { return aspect_; }
  float GetFovY() const { return fovy_; }
  void SetFovY(float f) { fovy_ = f; }
  std::array<float, 3> GetEye() const { return eye_; }
  std::array<float, 3> GetCenter() const { return center_; }
  std::array<float, 3> GetUp() const { return up_; }
  // GetCameraVectors (camera.cpp:336-341): forward = -dir, right = cross(up, fwd), up' = cross(fwd, right)
  void GetCameraVectors(std::array<float, 3>* fwd, std::array<float, 3>* up, std::array<float, 3>* right) const;
  bool Changing() const { return false; }
  cvr_camera ToC() const;

 private:
  std::array<float, 3> eye_{0, 0, 0}, center_{0, 0, 0}, up_{0, 1, 0};
  float fovy_ = 45.0f, aspect_ = 1.0f;
};

// CameraData / CameraStateList (vis_utils/camera.h, volvis_utils/camerastatelist.cpp:26-88):
// named camera states, each "name", "ARCBALL" (then eye, center, up as 9 numbers) or
// "FLIGHT" (no vectors), as data/#list_camera_states holds them.
struct CameraData {
  enum Type { FLIGHT = 0, ARCBALL = 1 };
  std::string cam_setup_name;
  int c_type = ARCBALL;
  std::array<float, 3> eye{0, 0, 0}, center{0, 0, 0}, up{0, 1, 0};
};

class CameraStateList {
 public:
  // False (with *err) when the file cannot be read or a record is malformed; a blank
  // trailing record (a final newline) is not a state.
  bool ReadCameraStates(const std::string& filepath, std::string* err = nullptr);
  int NumberOfCameraStates() const { return int(states_.size()); }
  const CameraData* GetCameraState(int i) const { return i >= 0 && i < NumberOfCameraStates() ? &states_[i] : nullptr; }
  int FindCameraState(const std::string& name) const;

 private:
  std::vector<CameraData> states_;
};

// LightSourceData / LightSourceList (volvis_utils/lightsourcelist.h, lightsourcelist.cpp:81-148):
// named lists of light sources, each light as position, forward (towards the light),
// up, right and the spot angle in degrees, as data/#list_light_sources holds them.  The
// light stores z_axis = -forward, as the reference does.
struct LightSourceData {
  std::array<float, 3> position{0, 0, 0};
  std::array<float, 3> z_axis{0, 0, 1}, y_axis{0, 1, 0}, x_axis{1, 0, 0};
  float spot_light_angle = 4.0f, spot_light_angle_rad = 4.0f * 3.14159265f / 180.0f;
  std::array<float, 3> Forward() const { return {-z_axis[0], -z_axis[1], -z_axis[2]}; }
};

struct LightSourceListItem {
  std::string l_name;
  std::vector<LightSourceData> m_lightsources;
};

class LightSourceList {
 public:
  bool ReadLightSourceLists(const std::string& filepath, std::string* err = nullptr);
  int NumberOfLists() const { return int(lists_.size()); }
  const LightSourceListItem* GetList(int i) const { return i >= 0 && i < NumberOfLists() ? &lists_[i] : nullptr; }
  int FindList(const std::string& name) const;

 private:
  std::vector<LightSourceListItem> lists_;
};

// volvis_utils/renderingparameters.cpp:17-31, 147-167 (headless: one light source).
class RenderingParameters {
 public:
  RenderingParameters();
  void SetScreenSize(int w, int h) { screen_width_ = w; screen_height_ = h; }
  int GetScreenWidth() const { return screen_width_; }
  int GetScreenHeight() const { return screen_height_; }
  void SetPhongParameters(float amb, float diff, float spec, float shini);
  float GetBlinnPhongKambient() const { return ka_; }
  float GetBlinnPhongKdiffuse() const { return kd_; }
  float GetBlinnPhongKspecular() const { return ks_; }
  float GetBlinnPhongNshininess() const { return n_; }
  void SetBlinnPhongLightingPosition(const std::array<float, 3>& p) { light_pos_ = p; }
  std::array<float, 3> GetBlinnPhongLightingPosition() const { return light_pos_; }
  void SetBlinnPhongLightSourceCameraVectors(const std::array<float, 3>& fwd, const std::array<float, 3>& up,
                                             const std::array<float, 3>& right);
  std::array<float, 3> GetBlinnPhongLightSourceCameraForward() const { return light_fwd_; }
  // CreateNewLightSource(LightSourceData) as the current light (renderingparameters.cpp:100-106)
  void SetLightSource(const LightSourceData& l);
  cvr_light ToLight() const;

 private:
  int screen_width_ = 768, screen_height_ = 768;
  float ka_ = 0.5f, kd_ = 0.5f, ks_ = 0.8f, n_ = 30.0f;
  std::array<float, 3> light_pos_{0, 0, 0}, light_fwd_{0, 0, 1}, light_up_{0, 1, 0}, light_right_{1, 0, 0};
};

// datamanager.h:85-99 (headless: the current structured volume and transfer function).
class DataManager {
 public:
  ~DataManager();
  void SetStructuredVolume(StructuredGridVolume* v);  // takes ownership
  void SetTransferFunction(TransferFunction1D* tf);   // takes ownership
  StructuredGridVolume* GetCurrentStructuredVolume() const { return vol_; }
  TransferFunction1D* GetCurrentTransferFunction() const { return tf_; }

 private:
  StructuredGridVolume* vol_ = nullptr;
  TransferFunction1D* tf_ = nullptr;
};

}  // namespace vis
