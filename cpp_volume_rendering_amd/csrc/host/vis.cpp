// Headless host data model (declarations: vis.h).
#include "vis.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace vis {

namespace {
void fail(std::string* err, const std::string& m) {
  if (err) *err = m;
}
std::array<float, 3> nrm(const std::array<double, 3>& v) {
  const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  return {float(v[0] / l), float(v[1] / l), float(v[2] / l)};
}
std::array<double, 3> crs(const std::array<double, 3>& a, const std::array<double, 3>& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
}  // namespace

// ------------------------------------------------------------------ volume
StructuredGridVolume::StructuredGridVolume(std::string name, int w, int h, int d)
    : name_(std::move(name)), w_(w), h_(h), d_(d) {}

void StructuredGridVolume::SetScale(double sx, double sy, double sz) { scale_ = {sx, sy, sz}; }

bool StructuredGridVolume::SetArrayData(std::vector<uint8_t> bytes, int bytes_per_value) {
  if ((bytes_per_value != 1 && bytes_per_value != 2) ||
      bytes.size() != size_t(w_) * h_ * d_ * size_t(bytes_per_value))
    return false;
  data_ = std::move(bytes);
  bpv_ = bytes_per_value;
  return true;
}

double StructuredGridVolume::GetNormalizedSample(int x, int y, int z) const {
  if (data_.empty() || x < 0 || y < 0 || z < 0 || x >= w_ || y >= h_ || z >= d_) return 0.0;
  const size_t i = size_t(x) + size_t(y) * w_ + size_t(z) * size_t(w_) * h_;
  if (bpv_ == 1) return double(data_[i]) / (256.0 - 1.0);
  const uint16_t* p = reinterpret_cast<const uint16_t*>(data_.data());
  return double(p[i]) / (65536.0 - 1.0);
}

double StructuredGridVolume::GetDiagonal() const {
  const double a = w_ * scale_[0], b = h_ * scale_[1], c = d_ * scale_[2];
  return std::sqrt(a * a + b * b + c * c);
}

// ------------------------------------------------------------------ transfer function
TransferFunction1D::TransferFunction1D(int max_value) : max_density_(max_value) {}

void TransferFunction1D::AddRGBControlPoint(double r, double g, double b, int isovalue) {
  rgb_iso_.insert(rgb_iso_.end(), {r, g, b, double(isovalue)});
  built_ = false;
}
void TransferFunction1D::AddAlphaControlPoint(double alpha, int isovalue) {
  alpha_iso_.insert(alpha_iso_.end(), {alpha, double(isovalue)});
  built_ = false;
}
void TransferFunction1D::ClearControlPoints() {
  rgb_iso_.clear();
  alpha_iso_.clear();
  built_ = false;
}

// Control point colours are float (TransferControlPoint holds a glm::vec4,
// transferfunction.cpp:8-24); the table is double and zero where no segment covers it.
bool TransferFunction1D::Build() {
  const int T = max_density_ + 1;
  table_.assign(size_t(T) * 4, 0.0);
  const int nr = int(rgb_iso_.size() / 4), na = int(alpha_iso_.size() / 2);
  for (int i = 0; i + 1 < nr; ++i) {
    const int i0 = int(rgb_iso_[i * 4 + 3]), i1 = int(rgb_iso_[i * 4 + 7]);
    if (i0 == i1 || i0 < 0 || i1 < 0 || i0 >= T || i1 >= T) return false;
    const float c0[3] = {float(rgb_iso_[i * 4]), float(rgb_iso_[i * 4 + 1]), float(rgb_iso_[i * 4 + 2])};
    const float c1[3] = {float(rgb_iso_[i * 4 + 4]), float(rgb_iso_[i * 4 + 5]), float(rgb_iso_[i * 4 + 6])};
    for (int x = i0; x <= i1; ++x) {
      const double k = double(x - i0) / double(i1 - i0);
      for (int c = 0; c < 3; ++c) table_[size_t(x) * 4 + c] = double(c0[c]) + double(c1[c] - c0[c]) * k;
    }
  }
  for (int i = 0; i + 1 < na; ++i) {
    const int i0 = int(alpha_iso_[i * 2 + 1]), i1 = int(alpha_iso_[i * 2 + 3]);
    if (i0 == i1 || i0 < 0 || i1 < 0 || i0 >= T || i1 >= T) return false;
    const float a0 = float(alpha_iso_[i * 2]), a1 = float(alpha_iso_[i * 2 + 2]);
    for (int x = i0; x <= i1; ++x) {
      const double k = double(x - i0) / double(i1 - i0);
      table_[size_t(x) * 4 + 3] = double(a0) + double(a1 - a0) * k;
    }
  }
  built_ = true;
  return true;
}

std::array<float, 4> TransferFunction1D::Get(double value, double max_data_value) {
  if (!built_) Build();
  if (max_data_value >= 0) value = value * (double(max_density_) / max_data_value);
  if (value < 0.0f || value > float(max_density_)) return {0.f, 0.f, 0.f, 0.f};
  std::array<float, 4> out;
  if (std::fabs(value - float(max_density_)) < 0.000001) {
    for (int c = 0; c < 4; ++c) out[c] = float(table_[size_t(max_density_) * 4 + c]);
    return out;
  }
  const int i = int(value);
  const double t = value - i;
  for (int c = 0; c < 4; ++c) out[c] = float((1.0 - t) * table_[size_t(i) * 4 + c] + t * table_[size_t(i + 1) * 4 + c]);
  return out;
}

float TransferFunction1D::GetExtN(double normalized_value) {
  const float val = Get(normalized_value, 1.0)[3];
  if (!extinction_coef_type_) return float(std::log(1.0 / (1.0 - double(val))));
  return val;
}

// ------------------------------------------------------------------ readers
TransferFunction1D* TransferFunctionReader::ReadTransferFunction(const std::string& file, std::string* err) {
  const size_t dot = file.find_last_of('.');
  if (dot != std::string::npos && file.substr(dot + 1) == "tf1d") return readtf1d(file, err);
  fail(err, "unsupported transfer function file: " + file);
  return nullptr;
}

// Line 1: interpolation (ignored); then type 0/1/2 (1: max density; 2: + extinction
// input flag); N rows "r g b isovalue"; M rows "a isovalue" (transferfunction1d.h:1-19).
TransferFunction1D* TransferFunctionReader::readtf1d(const std::string& file, std::string* err) {
  std::ifstream in(file);
  if (!in.is_open()) {
    fail(err, "cannot open " + file);
    return nullptr;
  }
  std::string interpolation;
  std::getline(in, interpolation);
  int init = -1;
  if (!(in >> init)) {
    fail(err, "missing type line in " + file);
    return nullptr;
  }
  int max_density = 255, extuse = 0;
  if (init == 2) {
    if (!(in >> max_density >> extuse)) { fail(err, "missing max density / extinction flag"); return nullptr; }
  } else if (init == 1) {
    if (!(in >> max_density)) { fail(err, "missing max density"); return nullptr; }
  }
  if (max_density < 1) { fail(err, "max density must be >= 1"); return nullptr; }
  TransferFunction1D* tf = new TransferFunction1D(max_density);
  tf->SetExtinctionCoefficientInput(extuse == 1);
  int n = 0;
  if (!(in >> n) || n < 0) { delete tf; fail(err, "bad rgb count"); return nullptr; }
  for (int i = 0; i < n; ++i) {
    double r, g, b;
    int iso;
    if (!(in >> r >> g >> b >> iso)) { delete tf; fail(err, "truncated rgb control points"); return nullptr; }
    tf->AddRGBControlPoint(r, g, b, iso);
  }
  int m = 0;
  if (!(in >> m) || m < 0) { delete tf; fail(err, "bad alpha count"); return nullptr; }
  for (int i = 0; i < m; ++i) {
    double a;
    int iso;
    if (!(in >> a >> iso)) { delete tf; fail(err, "truncated alpha control points"); return nullptr; }
    tf->AddAlphaControlPoint(a, iso);
  }
  if (!tf->Build()) { delete tf; fail(err, "control points with repeated or out-of-range isovalues"); return nullptr; }
  tf->SetName(file);
  return tf;
}

StructuredGridVolume* VolumeReader::ReadStructuredVolume(const std::string& file, std::string* err) {
  const size_t dot = file.find_last_of('.');
  if (dot != std::string::npos && file.substr(dot + 1) == "raw") return readraw(file, err);
  fail(err, "unsupported volume file (only .raw): " + file);
  return nullptr;
}

StructuredGridVolume* VolumeReader::readraw(const std::string& path, std::string* err) {
  // strip directories, then the .raw extension, then parse <bytes>.<W>x<H>x<D>
  std::string name = path.substr(path.find_last_of("/\\") == std::string::npos ? 0 : path.find_last_of("/\\") + 1);
  const size_t e = name.find_last_of('.');
  if (e == std::string::npos) { fail(err, "no extension"); return nullptr; }
  name = name.substr(0, e);
  const size_t s1 = name.find_last_of('.');
  if (s1 == std::string::npos) { fail(err, "expected <name>.<bytes>.<W>x<H>x<D>.raw"); return nullptr; }
  const std::string sizes = name.substr(s1 + 1);
  std::string rest = name.substr(0, s1);
  const size_t s2 = rest.find_last_of('.');
  if (s2 == std::string::npos) { fail(err, "expected <name>.<bytes>.<W>x<H>x<D>.raw"); return nullptr; }
  const int bpv = std::atoi(rest.substr(s2 + 1).c_str());
  int w = 0, h = 0, d = 0;
  if (std::sscanf(sizes.c_str(), "%dx%dx%d", &w, &h, &d) != 3 || w < 1 || h < 1 || d < 1 || (bpv != 1 && bpv != 2)) {
    fail(err, "bad sizes or bytes per value in " + path);
    return nullptr;
  }
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) { fail(err, "cannot open " + path); return nullptr; }
  std::vector<uint8_t> bytes(size_t(w) * h * d * size_t(bpv));
  in.read(reinterpret_cast<char*>(bytes.data()), std::streamsize(bytes.size()));
  if (size_t(in.gcount()) != bytes.size()) { fail(err, "short read: " + path); return nullptr; }
  StructuredGridVolume* v = new StructuredGridVolume(rest.substr(0, s2), w, h, d);
  v->SetScale(1.0, 1.0, 1.0);
  v->SetName(path);
  v->SetArrayData(std::move(bytes), bpv);
  return v;
}

// ------------------------------------------------------------------ camera / parameters
Camera::Camera() {}

void Camera::SetData(const std::array<float, 3>& eye, const std::array<float, 3>& center,
                     const std::array<float, 3>& up) {
  eye_ = eye;
  center_ = center;
  up_ = up;
}

void Camera::GetCameraVectors(std::array<float, 3>* fwd, std::array<float, 3>* up, std::array<float, 3>* right) const {
  const std::array<double, 3> f = {double(eye_[0]) - center_[0], double(eye_[1]) - center_[1], double(eye_[2]) - center_[2]};
  const std::array<float, 3> ff = nrm(f);
  const std::array<double, 3> fd = {ff[0], ff[1], ff[2]};
  const std::array<float, 3> rr = nrm(crs({double(up_[0]), double(up_[1]), double(up_[2])}, fd));
  const std::array<float, 3> uu = nrm(crs(fd, {double(rr[0]), double(rr[1]), double(rr[2])}));
  if (fwd) *fwd = ff;
  if (right) *right = rr;
  if (up) *up = uu;
}

cvr_camera Camera::ToC() const {
  cvr_camera c;
  for (int i = 0; i < 3; ++i) {
    c.eye[i] = eye_[i];
    c.center[i] = center_[i];
    c.up[i] = up_[i];
  }
  c.fovy_deg = fovy_;
  c.aspect = aspect_;
  return c;
}

RenderingParameters::RenderingParameters() {}

void RenderingParameters::SetPhongParameters(float amb, float diff, float spec, float shini) {
  ka_ = amb;
  kd_ = diff;
  ks_ = spec;
  n_ = shini;
}

void RenderingParameters::SetBlinnPhongLightSourceCameraVectors(const std::array<float, 3>& fwd,
                                                                const std::array<float, 3>& up,
                                                                const std::array<float, 3>& right) {
  light_fwd_ = fwd;
  light_up_ = up;
  light_right_ = right;
}

cvr_light RenderingParameters::ToLight() const {
  cvr_light l;
  for (int i = 0; i < 3; ++i) {
    l.position[i] = light_pos_[i];
    l.forward[i] = light_fwd_[i];
    l.up[i] = light_up_[i];
    l.right[i] = light_right_[i];
  }
  return l;
}

DataManager::~DataManager() {
  delete vol_;
  delete tf_;
}
void DataManager::SetStructuredVolume(StructuredGridVolume* v) {
  if (v != vol_) delete vol_;
  vol_ = v;
}
void DataManager::SetTransferFunction(TransferFunction1D* tf) {
  if (tf != tf_) delete tf_;
  tf_ = tf;
}

}  // namespace vis
