// GGEMSVolumeCreatorManager and its shapes (src/geometries/
// GGEMSVolumeCreatorManager.cc, GGEMSVolume.cc, GGEMSBox.cc, GGEMSTube.cc,
// GGEMSSphere.cc): a label volume painted on the device by the engine
// library's volume painter, written as MHD/raw plus a range file.
#pragma once

#include <map>
#include <string>

#include "../../../include/ggems_engine.h"

// global names: these are the opaque handle types of include/ggems.h
struct GGEMSVolumeCreatorManager {
  static GGEMSVolumeCreatorManager& get();
  float element_sizes[3] = {0.f, 0.f, 0.f};
  size_t dims[3] = {0, 0, 0};
  std::string data_type = "MET_FLOAT";
  std::string output_image, output_range;
  std::map<float, std::string> label_to_material;  // ordered by label, as the range file lists them
  ggh_volume* volume = nullptr;
  void set_element_sizes(float x, float y, float z, const std::string& unit);
  void set_dimensions(size_t x, size_t y, size_t z);
  void set_material(const std::string& material);  // label 0
  void set_data_type(const std::string& type);
  void add_label(float label, const std::string& material);
  void initialize();
  void draw(const ggh_volume_shape& shape);
  void write();
  void clean();
};

// GGEMSVolume and the three shapes: a label, a position and the shape's sizes
struct GGEMSVolume {
  int32_t shape = GGH_SHAPE_BOX;
  float label = 1.0f;
  float position[3] = {0.f, 0.f, 0.f};
  float size[3] = {0.f, 0.f, 0.f};
  void set_position(float x, float y, float z, const std::string& unit);
  void set_label(float v) { label = v; }
  void set_material(const std::string& material);
  void initialize();
  void draw();
};

// GGEMSBox(width, height, depth): width along x, height along y, depth along z
struct GGEMSBox : GGEMSVolume {
  GGEMSBox(float width, float height, float depth, const std::string& unit);
};
// GGEMSTube(radius_x, radius_y, height): an elliptic cylinder along z
struct GGEMSTube : GGEMSVolume {
  GGEMSTube(float radius_x, float radius_y, float height, const std::string& unit);
};
struct GGEMSSphere : GGEMSVolume {
  GGEMSSphere(float radius, const std::string& unit);
};
