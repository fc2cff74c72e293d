// GGEMSVolumeCreatorManager, GGEMSVolume and its shapes
// (src/geometries/GGEMSVolumeCreatorManager.cc:45-277, GGEMSVolume.cc:38-97).

#include "volume_creator.hh"

#include <fstream>
#include <vector>

#include <sstream>

#include "host_common.hh"
#include "simulation.hh"

using namespace gghost;

namespace {

int32_t met_code(const std::string& t) {
  static const char* names[] = {"MET_CHAR", "MET_UCHAR", "MET_SHORT", "MET_USHORT", "MET_INT", "MET_UINT",
                                "MET_FLOAT"};
  for (int i = 0; i < 7; ++i)
    if (t == names[i]) return i;
  return -1;
}

size_t met_bytes(int32_t code) { return code <= GGH_MET_UCHAR ? 1 : code <= GGH_MET_USHORT ? 2 : 4; }

void volume_check(int rc, const char* what) {
  if (rc < 0) raise("GGEMSVolumeCreatorManager", std::string(what) + ": " + ggh_volume_last_error());
}

}  // namespace

GGEMSVolumeCreatorManager& GGEMSVolumeCreatorManager::get() {
  static GGEMSVolumeCreatorManager m;
  return m;
}

void GGEMSVolumeCreatorManager::set_element_sizes(float x, float y, float z, const std::string& unit) {
  element_sizes[0] = distance_unit(x, unit);
  element_sizes[1] = distance_unit(y, unit);
  element_sizes[2] = distance_unit(z, unit);
}

void GGEMSVolumeCreatorManager::set_dimensions(size_t x, size_t y, size_t z) {
  dims[0] = x;
  dims[1] = y;
  dims[2] = z;
}

// SetMaterial: the background, label 0; an existing label 0 is kept (map insert)
void GGEMSVolumeCreatorManager::set_material(const std::string& material) {
  label_to_material.insert(std::make_pair(0.0f, material));
}

void GGEMSVolumeCreatorManager::set_data_type(const std::string& type) {
  if (met_code(type) < 0)
    raise("GGEMSVolumeCreatorManager",
          "Your type in not compatible. The type has to be: MET_CHAR, MET_UCHAR, MET_SHORT, MET_USHORT, MET_INT, "
          "MET_UINT or MET_FLOAT");
  data_type = type;
}

void GGEMSVolumeCreatorManager::add_label(float label, const std::string& material) {
  auto r = label_to_material.insert(std::make_pair(label, material));
  if (!r.second) {
    std::ostringstream oss;
    oss << "The label: " << r.first->first << " already exists...";
    raise("GGEMSVolumeCreatorManager", oss.str());
  }
}

// CheckParameters + AllocateImage (a zero-filled volume of the element type)
void GGEMSVolumeCreatorManager::initialize() {
  if (dims[0] == 0 && dims[1] == 0 && dims[2] == 0)
    raise("GGEMSVolumeCreatorManager", "Phantom dimensions have to be > 0!!!");
  if (element_sizes[0] == 0.0f && element_sizes[1] == 0.0f && element_sizes[2] == 0.0f)
    raise("GGEMSVolumeCreatorManager", "Phantom voxel sizes have to be > 0.0!!!");
  if (output_image.empty())
    raise("GGEMSVolumeCreatorManager", "A output image filename has to be done to phantom manager!!!");
  if (dims[0] > 0x7fffffff || dims[1] > 0x7fffffff || dims[2] > 0x7fffffff)
    raise("GGEMSVolumeCreatorManager", "Phantom dimensions are too large");
  if (volume) ggh_volume_destroy(volume);
  volume = nullptr;
  const int32_t d[3] = {(int32_t)dims[0], (int32_t)dims[1], (int32_t)dims[2]};
  volume_check(ggh_volume_create(GGEMSOpenCLManager::get().device, d, element_sizes, met_code(data_type), &volume),
               "allocate the volume");
}

void GGEMSVolumeCreatorManager::draw(const ggh_volume_shape& shape) {
  if (!volume) raise("GGEMSVolumeCreatorManager", "The volume is not initialized!!!");
  volume_check(ggh_volume_draw(volume, &shape), "draw");
}

// WriteMHDImage + WriteRangeToMaterialFile: "label label material" per line
void GGEMSVolumeCreatorManager::write() {
  if (!volume) raise("GGEMSVolumeCreatorManager", "The volume is not initialized!!!");
  const int32_t code = met_code(data_type);
  const size_t n = dims[0] * dims[1] * dims[2];
  std::vector<char> host(n * met_bytes(code));
  volume_check(ggh_volume_read(volume, host.data(), host.size()), "read the volume");
  const int d[3] = {(int)dims[0], (int)dims[1], (int)dims[2]};
  write_mhd(output_image, data_type.c_str(), d, element_sizes, host.data(), met_bytes(code));
  std::ofstream r(output_range, std::ios::out);
  for (const auto& kv : label_to_material) r << kv.first << " " << kv.first << " " << kv.second << std::endl;
}

void GGEMSVolumeCreatorManager::clean() {
  if (volume) ggh_volume_destroy(volume);
  volume = nullptr;
  for (int i = 0; i < 3; ++i) {
    element_sizes[i] = 0.0f;
    dims[i] = 0;
  }
  data_type = "MET_FLOAT";
  output_image.clear();
  output_range.clear();
  label_to_material.clear();
}

void GGEMSVolume::set_position(float x, float y, float z, const std::string& unit) {
  position[0] = distance_unit(x, unit);
  position[1] = distance_unit(y, unit);
  position[2] = distance_unit(z, unit);
}

void GGEMSVolume::set_material(const std::string& material) {
  GGEMSVolumeCreatorManager::get().add_label(label, material);
}

// the reference compiles its draw kernel for the volume's type here
void GGEMSVolume::initialize() {
  if (!GGEMSVolumeCreatorManager::get().volume)
    raise("GGEMSVolumeCreatorManager", "The volume is not initialized!!!");
}

void GGEMSVolume::draw() {
  ggh_volume_shape s;
  s.shape = shape;
  s.label = label;
  for (int i = 0; i < 3; ++i) {
    s.position[i] = position[i];
    s.size[i] = size[i];
  }
  GGEMSVolumeCreatorManager::get().draw(s);
}

// sizes in the order the draw kernels take them (DrawGGEMSBox.cl: height,
// width, depth; DrawGGEMSTube.cl: height, radius_x, radius_y)
GGEMSBox::GGEMSBox(float width, float height, float depth, const std::string& unit) {
  shape = GGH_SHAPE_BOX;
  size[0] = distance_unit(height, unit);
  size[1] = distance_unit(width, unit);
  size[2] = distance_unit(depth, unit);
}

GGEMSTube::GGEMSTube(float radius_x, float radius_y, float height, const std::string& unit) {
  shape = GGH_SHAPE_TUBE;
  size[0] = distance_unit(height, unit);
  size[1] = distance_unit(radius_x, unit);
  size[2] = distance_unit(radius_y, unit);
}

GGEMSSphere::GGEMSSphere(float radius, const std::string& unit) {
  shape = GGH_SHAPE_SPHERE;
  size[0] = distance_unit(radius, unit);
}
