// host_common.cc -- units, errors, physics-data blob, transforms, MHD I/O.
#include "host_common.hh"

#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

namespace gghost {

// ---------------------------------------------------------------------------
// units (GGEMSSystemOfUnits.hh: DistanceUnit / EnergyUnit / AngleUnit / DensityUnit)
// ---------------------------------------------------------------------------
float distance_unit(float v, const std::string& u) {
  if (u == "nm") return v * unit::nm;
  if (u == "um") return v * unit::um;
  if (u == "mm") return v * unit::mm;
  if (u == "cm") return v * unit::cm;
  if (u == "m") return v * unit::m;
  if (u == "km") return v * unit::km;
  if (u == "pc") return v * unit::pc;
  raise("DistanceUnit", "Unknown unit '" + u + "'!!! You have choice between: nm, um, mm, cm, m, km, pc");
}

float energy_unit(float v, const std::string& u) {
  if (u == "eV") return v * unit::eV;
  if (u == "keV") return v * unit::keV;
  if (u == "MeV") return v * unit::MeV;
  if (u == "GeV") return v * unit::GeV;
  if (u == "TeV") return v * unit::TeV;
  if (u == "PeV") return v * unit::PeV;
  raise("EnergyUnit", "Unknown unit '" + u + "'!!! You have choice between: eV, keV, MeV, GeV, TeV, PeV");
}

float angle_unit(float v, const std::string& u) {
  if (u == "rad") return v * unit::rad;
  if (u == "mrad") return v * unit::mrad;
  if (u == "deg") return v * unit::deg;
  if (u == "sr") return v * unit::sr;
  raise("AngleUnit", "Unknown unit '" + u + "'!!! You have choice between: rad, mrad, deg, sr");
}

float density_unit(float v, const std::string& u) {
  if (u == "g/cm3") return v * (unit::g / unit::cm3);
  if (u == "kg/m3") return v * (unit::kg / unit::m3);
  if (u == "mg/cm3") return v * (unit::mg / unit::cm3);
  raise("DensityUnit", "Unknown unit '" + u + "'!!! You have choice between: g/cm3, kg/m3, mg/cm3");
}

// ---------------------------------------------------------------------------
// errors and logging
// ---------------------------------------------------------------------------
namespace {
thread_local std::string g_last_error;
int g_verbosity = 0;
}  // namespace

void raise(const std::string& where, const std::string& msg) { throw Error("[GGEMS " + where + "] " + msg); }

void record_error(const std::string& msg) {
  g_last_error = msg;
  std::cerr << "[GGEMS error] " << msg << std::endl;
  const char* a = std::getenv("GGEMS_ABORT_ON_ERROR");
  if (a && a[0] == '1') std::abort();
}

const std::string& last_error() { return g_last_error; }
void clear_error() { g_last_error.clear(); }
int verbosity() { return g_verbosity; }
void set_verbosity(int level) { g_verbosity = level; }

void log(int level, const std::string& who, const std::string& msg) {
  if (level <= g_verbosity) std::cout << "[GGEMS " << who << "] " << msg << std::endl;
}

// ---------------------------------------------------------------------------
// physics data blob
// ---------------------------------------------------------------------------
namespace {
std::string library_dir() {
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&library_dir), &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    size_t s = p.find_last_of('/');
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
  }
  return ".";
}
}  // namespace

PhysicsData& PhysicsData::get() {
  static PhysicsData d;
  return d;
}

PhysicsData::PhysicsData() {
  const char* env = std::getenv("GGEMS_PHYSICS_DATA");
  path_ = env ? env : library_dir() + "/../data/_gen/ggems_physics.bin";
  std::ifstream f(path_, std::ios::binary);
  if (!f) raise("PhysicsData", "cannot open physics data " + path_ + " (build step extract_physics_data.py)");
  blob_.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  if (blob_.size() < 12 || std::memcmp(blob_.data(), "GGPD", 4) != 0) raise("PhysicsData", "bad physics data " + path_);
}

const void* PhysicsData::find(const std::string& name, uint32_t dtype, size_t* n) const {
  uint32_t count;
  std::memcpy(&count, blob_.data() + 8, 4);
  for (uint32_t i = 0; i < count; ++i) {
    const unsigned char* e = blob_.data() + 12 + (size_t)i * 72;
    if (std::strncmp(reinterpret_cast<const char*>(e), name.c_str(), 48) == 0) {
      uint32_t dt;
      uint64_t cnt, off;
      std::memcpy(&dt, e + 48, 4);
      std::memcpy(&cnt, e + 56, 8);
      std::memcpy(&off, e + 64, 8);
      if (dt != dtype) raise("PhysicsData", "type mismatch for " + name);
      if (n) *n = (size_t)cnt;
      return blob_.data() + off;
    }
  }
  raise("PhysicsData", "array " + name + " missing from " + path_);
}

const float* PhysicsData::f32(const std::string& name, size_t* n) const {
  return static_cast<const float*>(find(name, 0, n));
}
const int32_t* PhysicsData::i32(const std::string& name, size_t* n) const {
  return static_cast<const int32_t*>(find(name, 1, n));
}
const char* PhysicsData::bytes(const std::string& name, size_t* n) const {
  return static_cast<const char*>(find(name, 3, n));
}

// ---------------------------------------------------------------------------
// transforms
// ---------------------------------------------------------------------------
Mat4 mat_identity() { return Mat4{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}; }

Mat4 mat_mult(const Mat4& a, const Mat4& b) {
  Mat4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      r[i * 4 + j] = a[i * 4 + 0] * b[0 * 4 + j] + a[i * 4 + 1] * b[1 * 4 + j] + a[i * 4 + 2] * b[2 * 4 + j] +
                     a[i * 4 + 3] * b[3 * 4 + j];
  return r;
}

void Transform::set_axis(const float a[9]) {
  m = Mat4{a[0], a[1], a[2], 0, a[3], a[4], a[5], 0, a[6], a[7], a[8], 0, 0, 0, 0, 1};
}

void Transform::translate(float tx, float ty, float tz) {
  Mat4 t{1, 0, 0, tx, 0, 1, 0, ty, 0, 0, 1, tz, 0, 0, 0, 1};
  m = mat_mult(t, m);
}

void Transform::rotate(float rx, float ry, float rz) {
  float c = std::cos(rx), s = std::sin(rx);
  Mat4 Rx{1, 0, 0, 0, 0, c, -s, 0, 0, s, c, 0, 0, 0, 0, 1};
  c = std::cos(ry);
  s = std::sin(ry);
  Mat4 Ry{c, 0, s, 0, 0, 1, 0, 0, -s, 0, c, 0, 0, 0, 0, 1};
  c = std::cos(rz);
  s = std::sin(rz);
  Mat4 Rz{c, -s, 0, 0, s, c, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  Mat4 r = mat_mult(Rz, mat_mult(Ry, Rx));
  m = mat_mult(r, m);
}

// ---------------------------------------------------------------------------
// MHD
// ---------------------------------------------------------------------------
MHDHeader read_mhd_header(const std::string& path) {
  std::ifstream in(path);
  if (!in) raise("GGEMSMHDImage", "Problem reading filename: " + path);
  MHDHeader h;
  std::string dir;
  size_t s = path.find_last_of("/\\");
  if (s != std::string::npos) dir = path.substr(0, s + 1);
  std::string line, raw;
  while (std::getline(in, line)) {
    size_t first = line.find_first_not_of("\t ");
    if (first == std::string::npos || line[first] == '#') continue;
    size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    std::string key;
    for (char ch : line.substr(first, eq - first))
      if (ch != ' ' && ch != '\t' && ch != '\r') key += ch;
    std::istringstream iss(line.substr(eq + 1));
    if (key == "DimSize") iss >> h.dim[0] >> h.dim[1] >> h.dim[2];
    else if (key == "ElementSpacing") iss >> h.spacing[0] >> h.spacing[1] >> h.spacing[2];
    else if (key == "ElementType") iss >> h.type;
    else if (key == "ElementDataFile") iss >> raw;
  }
  if (h.dim[0] <= 0 || h.dim[1] <= 0 || h.dim[2] <= 0)
    raise("GGEMSMHDImage", "Dimension invalid for the key 'DimSize'!!! The values have to be > 0");
  if (h.spacing[0] == 0.0f || h.spacing[1] == 0.0f || h.spacing[2] == 0.0f)
    raise("GGEMSMHDImage", "Voxel size invalid for the key 'ElementSpacing'!!! The values have to be > 0");
  if (h.type.empty()) raise("GGEMSMHDImage", "Value invalid for the key 'ElementType'!!!");
  if (raw.empty()) raise("GGEMSMHDImage", "Value invalid for the key 'ElementDataFile'!!! A filename for raw data has to be given");
  h.raw_file = dir + raw;
  return h;
}

// GGEMSMHDImage::SetOutputFileName + Write<T> (src/io/GGEMSMHDImage.cc:67-88,
// include/GGEMS/io/GGEMSMHDImage.hh:195-221)
void write_mhd(const std::string& filename, const char* type, const int dim[3], const float spacing[3],
               const void* data, size_t element_size) {
  std::string header, dir, raw;
  size_t found_mhd = filename.find(".mhd");
  header = found_mhd == std::string::npos ? filename + ".mhd" : filename;
  size_t found_dir = filename.find_last_of("/\\");
  if (found_dir != std::string::npos) {
    dir = filename.substr(0, found_dir + 1);
    raw = filename.substr(found_dir + 1, found_mhd - found_dir - 1) + ".raw";
  } else {
    raw = filename.substr(0, found_mhd) + ".raw";
  }
  std::ofstream h(header);
  if (!h) raise("GGEMSMHDImage", "cannot write " + header);
  h << "ObjectType = Image" << std::endl;
  h << "BinaryDataByteOrderMSB = False" << std::endl;
  h << "NDims = 3" << std::endl;
  h << "ElementSpacing = " << spacing[0] << " " << spacing[1] << " " << spacing[2] << std::endl;
  h << "DimSize = " << dim[0] << " " << dim[1] << " " << dim[2] << std::endl;
  h << "ElementType = " << type << std::endl;
  h << "ElementDataFile = " << raw << std::endl;
  std::ofstream r(dir + raw, std::ios::binary);
  if (!r) raise("GGEMSMHDImage", "cannot write " + dir + raw);
  r.write(static_cast<const char*>(data), (std::streamsize)((size_t)dim[0] * dim[1] * dim[2] * element_size));
}

}  // namespace gghost
