// host_common.hh -- shared pieces of the GGEMS-compatible host library:
// unit strings, error reporting, the physics-data blob, FP32 4x4 transforms.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace gghost {

// ---- units (include/GGEMS/tools/GGEMSSystemOfUnits.hh:54-192) -------------
namespace unit {
constexpr float mm = 1.0f, cm = 10.f, m = 1.e3f, km = 1.e6f, um = 1.e-3f, nm = 1.e-6f, pc = 3.0856775807e+19f;
constexpr float eV = 1.e-6f, keV = 1.e-3f, MeV = 1.f, GeV = 1.e+3f, TeV = 1.e+6f, PeV = 1.e+9f;
constexpr float rad = 1.0f, mrad = 1.e-3f, deg = 3.141592653589793238463f / 180.0f, sr = 1.0f;
constexpr float kg = 6.241509704e+24f, g = 6.241509704e+21f, mg = 6.241509704e+18f;
constexpr float cm3 = 1.e3f, m3 = 1.e9f, mol = 1.0f;
constexpr float barn = 1.e-22f;
constexpr float Gy = 1.0e-12f;
}  // namespace unit

float distance_unit(float v, const std::string& u);
float energy_unit(float v, const std::string& u);
float angle_unit(float v, const std::string& u);
float density_unit(float v, const std::string& u);

// ---- errors ------------------------------------------------------------------
// GGEMS reports a failed check by printing the message and throwing
// (src/tools/GGEMSTools.cc:63-69).  Here the C API catches it at the
// boundary, prints it, records it for get_last_error_ggems() and returns;
// GGEMS_ABORT_ON_ERROR=1 restores GGEMS's abort-the-process behaviour.
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
[[noreturn]] void raise(const std::string& where, const std::string& msg);
void record_error(const std::string& msg);
const std::string& last_error();
void clear_error();

// verbosity (src/tools/GGEMSPrint.cc)
int verbosity();
void set_verbosity(int level);
void log(int level, const std::string& who, const std::string& msg);

// ---- physics data blob ----------------------------------------------------------
// Arrays extracted from GGEMS's data tables at build time
// (ggems_amd/data/extract_physics_data.py).
class PhysicsData {
 public:
  static PhysicsData& get();
  const float* f32(const std::string& name, size_t* n = nullptr) const;
  const int32_t* i32(const std::string& name, size_t* n = nullptr) const;
  const char* bytes(const std::string& name, size_t* n = nullptr) const;
  const std::string& path() const { return path_; }

 private:
  PhysicsData();
  const void* find(const std::string& name, uint32_t dtype, size_t* n) const;
  std::vector<unsigned char> blob_;
  std::string path_;
};

// ---- transforms (src/maths/GGEMSGeometryTransformation.cc:40-336) --------------
using Mat4 = std::array<float, 16>;  // row major
Mat4 mat_identity();
Mat4 mat_mult(const Mat4& a, const Mat4& b);  // GGfloat44MultGGfloat44, FP32 left to right

struct Transform {
  Mat4 m = mat_identity();
  void set_axis(const float axis[9]);  // resets to the local axis matrix
  void translate(float tx, float ty, float tz);  // left-multiplies T
  void rotate(float rx, float ry, float rz);     // left-multiplies Rz*Ry*Rx
};

// ---- MHD images (src/io/GGEMSMHDImage.cc) ----------------------------------------
struct MHDHeader {
  int dim[3] = {0, 0, 0};
  float spacing[3] = {0.f, 0.f, 0.f};
  std::string type;
  std::string raw_file;  // path resolved against the header directory
};
MHDHeader read_mhd_header(const std::string& path);
// writes <base>.mhd + raw; `type` is the MET_* name, data holds dim product elements
void write_mhd(const std::string& basename, const char* type, const int dim[3], const float spacing[3],
               const void* data, size_t element_size);

}  // namespace gghost
