// Headless host data model (declarations: vis.h).
#include "vis.h"

#include <algorithm>
#include <cctype>
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

namespace {
std::string lower(std::string t) {
  for (char& c : t) c = char(std::tolower(static_cast<unsigned char>(c)));
  return t;
}
std::string trim(const std::string& t) {
  const size_t a = t.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  return t.substr(a, t.find_last_not_of(" \t\r\n") - a + 1);
}
// folder of a path (no trailing separator; "." for a bare name) and its file name
std::pair<std::string, std::string> split_path(const std::string& p) {
  const size_t t = p.find_last_of("/\\");
  if (t == std::string::npos) return {".", p};
  return {p.substr(0, t), p.substr(t + 1)};
}
// "key: value" lines up to the first empty line (or the end); '#' lines are comments
std::vector<std::pair<std::string, std::string>> header_keys(std::istream& in, size_t* data_offset) {
  std::vector<std::pair<std::string, std::string>> kv;
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (trim(line).empty()) break;
    if (line[0] == '#') continue;
    const size_t c = line.find(':');
    if (c == std::string::npos) continue;
    kv.emplace_back(trim(line.substr(0, c)), trim(line.substr(c + 1)));
  }
  if (data_offset) {
    const std::streampos p = in.tellg();
    *data_offset = p < 0 ? 0 : size_t(p);
  }
  return kv;
}
}  // namespace

StructuredGridVolume* VolumeReader::ReadStructuredVolume(const std::string& file, std::string* err) {
  const std::string ext = lower(DataProvider::GetExtension(file));
  if (ext == "raw") return readraw(file, err);
  if (ext == "syn") return readsyn(file, err);
  if (ext == "nrrd") return readnrrd(file, err);
  if (ext == "nhrd" || ext == "nhdr") return readnhrd(file, err);
  if (ext == "dat") return readdat(file, err);
  if (ext == "pvm" || ext == "pvmold") {
    fail(err, "the .pvm format (third-party compressed volumes) is not supported: " + file);
    return nullptr;
  }
  fail(err, "unsupported volume file (.raw, .syn, .nrrd, .nhrd, .dat): " + file);
  return nullptr;
}

bool VolumeReader::SetArrayDataFromRawFile(const std::string& path, StructuredGridVolume* sg, int bpv,
                                           std::string* err, size_t offset) {
  if (bpv != 1 && bpv != 2) { fail(err, "bytes per value must be 1 or 2"); return false; }
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) { fail(err, "cannot open " + path); return false; }
  if (offset) in.seekg(std::streamoff(offset));
  std::vector<uint8_t> bytes(size_t(sg->GetWidth()) * sg->GetHeight() * sg->GetDepth() * size_t(bpv));
  in.read(reinterpret_cast<char*>(bytes.data()), std::streamsize(bytes.size()));
  if (size_t(in.gcount()) != bytes.size()) { fail(err, "short read: " + path); return false; }
  return sg->SetArrayData(std::move(bytes), bpv);
}

StructuredGridVolume* VolumeReader::readraw(const std::string& path, std::string* err) {
  // strip directories, then the .raw extension, then parse <bytes>.<W>x<H>x<D>
  std::string name = split_path(path).second;
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
  std::unique_ptr<StructuredGridVolume> v(new StructuredGridVolume(rest.substr(0, s2), w, h, d));
  v->SetScale(1.0, 1.0, 1.0);
  v->SetName(path);
  if (!SetArrayDataFromRawFile(path, v.get(), bpv, err)) return nullptr;
  return v.release();
}

StructuredGridVolume* VolumeReader::readsyn(const std::string& path, std::string* err) {
  std::ifstream in(path);
  if (!in.is_open()) { fail(err, "cannot open " + path); return nullptr; }
  long long w = 0, h = 0, d = 0;
  if (!(in >> w >> h >> d) || w < 1 || h < 1 || d < 1 || w * h * d > (1ll << 34)) {
    fail(err, "bad .syn size line in " + path);
    return nullptr;
  }
  // voxels the file does not set stay 0
  std::vector<uint8_t> bytes(size_t(w * h * d), 0);
  auto in_box = [&](long long x, long long y, long long z) { return x >= 0 && y >= 0 && z >= 0 && x < w && y < h && z < d; };
  long long kind = 0;
  while (in >> kind) {
    if (kind == 1) {
      long long x0, y0, z0, x1, y1, z1, v;
      if (!(in >> x0 >> y0 >> z0 >> x1 >> y1 >> z1 >> v)) { fail(err, "truncated box record in " + path); return nullptr; }
      if (x0 < x1 && y0 < y1 && z0 < z1 && !(in_box(x0, y0, z0) && in_box(x1 - 1, y1 - 1, z1 - 1))) {
        fail(err, "box outside the volume in " + path);
        return nullptr;
      }
      for (long long z = z0; z < z1; ++z)
        for (long long y = y0; y < y1; ++y)
          for (long long x = x0; x < x1; ++x) bytes[size_t(x + w * y + w * h * z)] = uint8_t(v);
    } else {
      long long x, y, z, v;
      if (!(in >> x >> y >> z >> v)) { fail(err, "truncated voxel record in " + path); return nullptr; }
      if (!in_box(x, y, z)) { fail(err, "voxel outside the volume in " + path); return nullptr; }
      bytes[size_t(x + w * y + w * h * z)] = uint8_t(v);
    }
  }
  if (!in.eof()) { fail(err, "unreadable record in " + path); return nullptr; }
  StructuredGridVolume* sg = new StructuredGridVolume(path, int(w), int(h), int(d));
  sg->SetScale(1.0, 1.0, 1.0);
  sg->SetArrayData(std::move(bytes), 1);
  return sg;
}

StructuredGridVolume* VolumeReader::readnrrd(const std::string& path, std::string* err) {
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) { fail(err, "cannot open " + path); return nullptr; }
  std::string magic;
  std::getline(in, magic);
  if (magic.compare(0, 4, "NRRD") != 0) { fail(err, "not a NRRD header (first line NRRDxxxx): " + path); return nullptr; }
  size_t data_offset = 0;
  const auto kv = header_keys(in, &data_offset);
  std::string type, sizes, data_file, encoding = "raw", endian = "little";
  int dimension = 3;
  for (const auto& p : kv) {
    const std::string k = lower(p.first);
    if (k == "type") type = lower(p.second);
    else if (k == "sizes") sizes = p.second;
    else if (k == "dimension") dimension = std::atoi(p.second.c_str());
    else if (k == "data file" || k == "datafile") data_file = p.second;
    else if (k == "encoding") encoding = lower(p.second);
    else if (k == "endian") endian = lower(p.second);
  }
  int res[3] = {0, 0, 0};
  if (dimension != 3 || std::sscanf(sizes.c_str(), "%d %d %d", &res[0], &res[1], &res[2]) != 3 || res[0] < 1 ||
      res[1] < 1 || res[2] < 1) {
    fail(err, "NRRD: need dimension 3 and three sizes in " + path);
    return nullptr;
  }
  int bpv = 0;
  if (type == "uchar" || type == "unsigned char" || type == "uint8" || type == "uint8_t") bpv = 1;
  if (type == "ushort" || type == "unsigned short" || type == "uint16" || type == "uint16_t" ||
      type == "unsigned short int")
    bpv = 2;
  if (!bpv) { fail(err, "NRRD: type must be an unsigned 8 or 16 bit integer, got '" + type + "'"); return nullptr; }
  if (encoding != "raw") { fail(err, "NRRD: only raw encoding is supported, got '" + encoding + "'"); return nullptr; }
  if (bpv == 2 && endian != "little") { fail(err, "NRRD: only little-endian 16 bit data is supported"); return nullptr; }
  const auto dn = split_path(path);
  std::unique_ptr<StructuredGridVolume> sg(new StructuredGridVolume(dn.second, res[0], res[1], res[2]));
  sg->SetScale(1.0, 1.0, 1.0);
  // a detached data file is relative to the header's folder; without one the data
  // follows the header's empty line in the same file
  const bool ok = data_file.empty() ? SetArrayDataFromRawFile(path, sg.get(), bpv, err, data_offset)
                                    : SetArrayDataFromRawFile(dn.first + "/" + data_file, sg.get(), bpv, err);
  if (!ok) return nullptr;
  return sg.release();
}

StructuredGridVolume* VolumeReader::readnhrd(const std::string& path, std::string* err) { return readnrrd(path, err); }

StructuredGridVolume* VolumeReader::readdat(const std::string& path, std::string* err) {
  std::ifstream in(path);
  if (!in.is_open()) { fail(err, "cannot open " + path); return nullptr; }
  const auto kv = header_keys(in, nullptr);
  std::string object, format;
  int res[3] = {0, 0, 0};
  double th[3] = {1.0, 1.0, 1.0};
  bool have_res = false;
  for (const auto& p : kv) {
    if (p.first == "ObjectFileName") object = p.second;
    else if (p.first == "Resolution") have_res = std::sscanf(p.second.c_str(), "%d %d %d", &res[0], &res[1], &res[2]) == 3;
    else if (p.first == "SliceThickness") {
      if (std::sscanf(p.second.c_str(), "%lf %lf %lf", &th[0], &th[1], &th[2]) != 3) {
        fail(err, ".dat: SliceThickness needs three values in " + path);
        return nullptr;
      }
    } else if (p.first == "Format") format = p.second;
  }
  if (!have_res || res[0] < 1 || res[1] < 1 || res[2] < 1 || object.empty()) {
    fail(err, ".dat: need ObjectFileName and Resolution in " + path);
    return nullptr;
  }
  if (!(th[0] > 0 && th[1] > 0 && th[2] > 0)) { fail(err, ".dat: SliceThickness must be positive"); return nullptr; }
  int bpv = 0;
  if (format.find("UCHAR") != std::string::npos) bpv = 1;
  else if (format.find("USHORT") != std::string::npos) bpv = 2;
  if (!bpv) { fail(err, ".dat: Format must be UCHAR or USHORT in " + path); return nullptr; }
  const auto dn = split_path(path);
  std::unique_ptr<StructuredGridVolume> sg(new StructuredGridVolume(dn.second, res[0], res[1], res[2]));
  sg->SetScale(th[0], th[1], th[2]);
  if (!SetArrayDataFromRawFile(dn.first + "/" + object, sg.get(), bpv, err)) return nullptr;
  return sg.release();
}

// ------------------------------------------------------------------ dataset lists
void DataProvider::ClearStructuredGridFileList() {
  names_.clear();
  paths_.clear();
  rows_.clear();
  tags_.clear();
}

std::string DataProvider::GetExtension(const std::string& filename) {
  const size_t slash = filename.find_last_of("/\\"), dot = filename.find_last_of('.');
  if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "";
  return filename.substr(dot + 1);
}

int DataProvider::Split(std::vector<std::string>& tags, const std::string& line, const std::string& separator) {
  tags.clear();
  size_t a = 0;
  for (;;) {
    const size_t b = line.find(separator, a);
    if (b == std::string::npos) break;
    tags.push_back(line.substr(a, b - a));
    a = b + separator.size();
  }
  tags.push_back(line.substr(a));
  return int(tags.size());
}

int DataProvider::GetBytesPerValueByType(const std::string& t) {
  if (t == "uchar" || t == "uint8") return 1;
  if (t == "ushort" || t == "uint16") return 2;
  return 0;
}

bool DataProvider::SetStructuredGridFileList(const std::string& file_path, const std::string& file_name,
                                             std::string* err) {
  ClearStructuredGridFileList();
  dir_ = file_path.empty() ? std::string(".") : file_path;
  while (dir_.size() > 1 && (dir_.back() == '/' || dir_.back() == '\\')) dir_.pop_back();
  list_ = file_name;
  std::ifstream in(dir_ + "/" + list_);
  if (!in.is_open()) { fail(err, "cannot open the dataset list " + dir_ + "/" + list_); return false; }
  csv_ = file_name.find('.') != std::string::npos;
  std::string line;
  if (!csv_) {
    // "<relative path>   <ui name>" per line; an empty name falls back to the path
    while (std::getline(in, line)) {
      if (!line.empty() && line.back() == '\r') line.pop_back();
      const size_t p0 = line.find('<'), p1 = line.find('>');
      if (p0 == std::string::npos || p1 == std::string::npos || p1 < p0) continue;
      const size_t n0 = line.find_last_of('<'), n1 = line.find_last_of('>');
      const std::string path = line.substr(p0 + 1, p1 - p0 - 1);
      const std::string name = n0 > p0 ? line.substr(n0 + 1, n1 - n0 - 1) : std::string();
      paths_.push_back(path);
      names_.push_back(name.empty() ? path : name);
    }
    return true;
  }
  if (lower(GetExtension(file_name)) != "csv") { fail(err, "dataset lists are CSV files or extension-less lists"); return false; }
  if (!std::getline(in, line)) { fail(err, "empty dataset list"); return false; }
  if (!line.empty() && line.back() == '\r') line.pop_back();
  sep_ = ";";
  if (line.compare(0, 3, "sep") == 0) {
    sep_ = line.substr(line.size() - 1, 1);
    if (!std::getline(in, line)) { fail(err, "dataset list without a header"); return false; }
    if (!line.empty() && line.back() == '\r') line.pop_back();
  }
  Split(tags_, line, sep_);
  for (const char* need : {"Name", "Source", "File", "Size", "Spacing", "Type"})
    if (std::find(tags_.begin(), tags_.end(), need) == tags_.end()) {
      fail(err, std::string("dataset list lacks the column ") + need);
      return false;
    }
  const int iname = int(std::find(tags_.begin(), tags_.end(), "Name") - tags_.begin());
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (trim(line).empty()) continue;
    std::vector<std::string> v;
    if (Split(v, line, sep_) != int(tags_.size())) { fail(err, "dataset list row with a wrong column count: " + line); return false; }
    names_.push_back(v[iname]);
    rows_.push_back(std::move(v));
  }
  return true;
}

int DataProvider::FindStructuredGridName(const std::string& name) const {
  for (size_t i = 0; i < names_.size(); ++i)
    if (names_[i] == name) return int(i);
  return -1;
}

StructuredGridVolume* DataProvider::LoadStructuredGridFromRawFile(const std::string& name, const std::string& raw_file_path,
                                                                 const std::array<int, 3>& r,
                                                                 const std::array<double, 3>& scale, int bpv,
                                                                 std::string* err) {
  if (r[0] < 1 || r[1] < 1 || r[2] < 1 || !(scale[0] > 0 && scale[1] > 0 && scale[2] > 0)) {
    fail(err, "bad size or spacing for " + name);
    return nullptr;
  }
  std::unique_ptr<StructuredGridVolume> sg(new StructuredGridVolume(name, r[0], r[1], r[2]));
  sg->SetScale(scale[0], scale[1], scale[2]);
  if (!VolumeReader::SetArrayDataFromRawFile(raw_file_path, sg.get(), bpv, err)) return nullptr;
  return sg.release();
}

StructuredGridVolume* DataProvider::LoadStructuredGrid(unsigned int id, std::string* err) {
  if (id >= names_.size()) { fail(err, "no dataset " + std::to_string(id)); return nullptr; }
  if (!csv_) return VolumeReader::ReadStructuredVolume(dir_ + "/" + paths_[id], err);
  const std::vector<std::string>& v = rows_[id];
  auto col = [&](const char* t) { return v[std::find(tags_.begin(), tags_.end(), t) - tags_.begin()]; };
  const std::string file = dir_ + "/" + col("Source") + "/" + col("File");
  if (lower(GetExtension(file)) != "raw") return VolumeReader::ReadStructuredVolume(file, err);
  std::array<int, 3> r{0, 0, 0};
  std::array<double, 3> sp{0, 0, 0};
  if (std::sscanf(col("Size").c_str(), "%dx%dx%d", &r[0], &r[1], &r[2]) != 3 ||
      std::sscanf(col("Spacing").c_str(), "%lfx%lfx%lf", &sp[0], &sp[1], &sp[2]) != 3) {
    fail(err, "dataset " + names_[id] + ": Size must be WxHxD and Spacing sxxsyxsz");
    return nullptr;
  }
  const int bpv = GetBytesPerValueByType(col("Type"));
  if (!bpv) { fail(err, "dataset " + names_[id] + ": Type must be uchar, uint8, ushort or uint16"); return nullptr; }
  return LoadStructuredGridFromRawFile(names_[id], file, r, sp, bpv, err);
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

void RenderingParameters::SetLightSource(const LightSourceData& l) {
  light_pos_ = l.position;
  light_fwd_ = l.Forward();
  light_up_ = l.y_axis;
  light_right_ = l.x_axis;
}

// ------------------------------------------------------------------ camera states, light lists
namespace {
// Whitespace-separated numbers from a stream; false when fewer than n could be read.
bool read_floats(std::istream& in, float* v, int n) {
  for (int i = 0; i < n; ++i)
    if (!(in >> v[i])) return false;
  return true;
}
// Next line that is not blank (false at the end of the stream).
bool next_nonblank(std::istream& in, std::string& line) {
  while (std::getline(in, line))
    if (!trim(line).empty()) return true;
  return false;
}
bool list_error(std::string* err, const std::string& msg) {
  fail(err, msg);
  return false;
}
}  // namespace

bool CameraStateList::ReadCameraStates(const std::string& filepath, std::string* err) {
  states_.clear();
  std::ifstream f(filepath);
  if (!f) return list_error(err, "cannot open camera states " + filepath);
  std::string name, type;
  while (next_nonblank(f, name)) {
    CameraData c;
    c.cam_setup_name = trim(name);
    if (!std::getline(f, type)) return list_error(err, "camera state '" + c.cam_setup_name + "': missing type line");
    type = trim(type);
    if (type == "FLIGHT") {
      c.c_type = CameraData::FLIGHT;
    } else if (type == "ARCBALL") {
      c.c_type = CameraData::ARCBALL;
      float v[9];
      if (!read_floats(f, v, 9)) return list_error(err, "camera state '" + c.cam_setup_name + "': needs eye, center, up");
      c.eye = {v[0], v[1], v[2]};
      c.center = {v[3], v[4], v[5]};
      c.up = {v[6], v[7], v[8]};
      std::string rest;
      std::getline(f, rest);  // the end of the up line
    } else {
      return list_error(err, "camera state '" + c.cam_setup_name + "': type must be ARCBALL or FLIGHT, not '" + type + "'");
    }
    states_.push_back(c);
  }
  if (states_.empty()) return list_error(err, "no camera states in " + filepath);
  return true;
}

int CameraStateList::FindCameraState(const std::string& name) const {
  for (int i = 0; i < NumberOfCameraStates(); ++i)
    if (states_[i].cam_setup_name == name) return i;
  return -1;
}

bool LightSourceList::ReadLightSourceLists(const std::string& filepath, std::string* err) {
  lists_.clear();
  std::ifstream f(filepath);
  if (!f) return list_error(err, "cannot open light source lists " + filepath);
  std::string name;
  while (next_nonblank(f, name)) {
    LightSourceListItem item;
    item.l_name = trim(name);
    int n = 0;
    if (!(f >> n) || n < 0 || n > 64) return list_error(err, "light source list '" + item.l_name + "': bad light count");
    for (int i = 0; i < n; ++i) {
      float v[13];
      if (!read_floats(f, v, 13))
        return list_error(err, "light source list '" + item.l_name + "': needs position, forward, up, right, spot angle");
      LightSourceData l;
      l.position = {v[0], v[1], v[2]};
      l.z_axis = {-v[3], -v[4], -v[5]};
      l.y_axis = {v[6], v[7], v[8]};
      l.x_axis = {v[9], v[10], v[11]};
      l.spot_light_angle = v[12];
      l.spot_light_angle_rad = (v[12] * 3.14159265358979f) / 180.0f;
      item.m_lightsources.push_back(l);
    }
    std::string rest;
    std::getline(f, rest);  // the end of the spot-angle line
    lists_.push_back(item);
  }
  if (lists_.empty()) return list_error(err, "no light source lists in " + filepath);
  return true;
}

int LightSourceList::FindList(const std::string& name) const {
  for (int i = 0; i < NumberOfLists(); ++i)
    if (lists_[i].l_name == name) return i;
  return -1;
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
