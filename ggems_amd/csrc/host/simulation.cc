// simulation.cc -- GGEMS-compatible host objects and the run orchestration.
#include "simulation.hh"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>

using namespace gghost;

namespace {
void engine_check(int rc, const char* what) {
  if (rc < 0) raise("GGEMSEngine", std::string(what) + ": " + ggh_last_error());
}
ggh_context* g_current_ctx = nullptr;
}  // namespace

ggh_context* current_engine_context() { return g_current_ctx; }

// ---------------------------------------------------------------------------
// singletons
// ---------------------------------------------------------------------------
GGEMSProcessesManager& GGEMSProcessesManager::get() {
  static GGEMSProcessesManager m;
  return m;
}
GGEMSRangeCutsManager& GGEMSRangeCutsManager::get() {
  static GGEMSRangeCutsManager m;
  return m;
}
GGEMSMaterialsDatabaseManager& GGEMSMaterialsDatabaseManager::get() {
  static GGEMSMaterialsDatabaseManager m;
  return m;
}
GGEMSSourceManager& GGEMSSourceManager::get() {
  static GGEMSSourceManager m;
  return m;
}
GGEMSNavigatorManager& GGEMSNavigatorManager::get() {
  static GGEMSNavigatorManager m;
  return m;
}
GGEMSOpenCLManager& GGEMSOpenCLManager::get() {
  static GGEMSOpenCLManager m;
  return m;
}
GGEMSProfilerManager& GGEMSProfilerManager::get() {
  static GGEMSProfilerManager m;
  return m;
}
GGEMSRAMManager& GGEMSRAMManager::get() {
  static GGEMSRAMManager m;
  return m;
}

GGEMSNavigator* GGEMSNavigatorManager::find(const std::string& name) const {
  for (GGEMSNavigator* n : navigators)
    if (n->name == name) return n;
  raise("GGEMSNavigatorManager", "Navigator '" + name + "' not found!!!");
}

// The selection strings of GGEMSOpenCLManager::DeviceToActivate
// (src/global/GGEMSOpenCLManager.cc:716-766) map onto HIP devices: "gpu",
// "all" and vendor filters select device 0; "i;j" selects the listed indices.
// One device is transported on; further devices are reported and ignored.
void GGEMSOpenCLManager::activate(const std::string& type, const std::string& vendor) {
  (void)vendor;
  int count = 0;
  if (ggh_device_count(&count) < 0 || count == 0) raise("GGEMSOpenCLManager", "no HIP device available");
  if (type == "cpu") raise("GGEMSOpenCLManager", "no CPU device: the transport runs on HIP GPUs only");
  std::vector<int> ids;
  if (type == "gpu" || type == "all" || type.empty()) {
    ids.push_back(0);
    if (type == "all" && count > 1) log(0, "GGEMSOpenCLManager", "several devices present; transporting on device 0");
  } else {
    std::stringstream ss(type);
    std::string tok;
    while (std::getline(ss, tok, ';')) {
      if (tok.empty()) continue;
      int id = std::atoi(tok.c_str());
      if (id < 0 || id >= count) raise("GGEMSOpenCLManager", "device index " + tok + " out of range");
      ids.push_back(id);
    }
    if (ids.empty()) raise("GGEMSOpenCLManager", "unknown device selection '" + type + "'");
    if (ids.size() > 1) log(0, "GGEMSOpenCLManager", "several devices requested; transporting on the first one");
  }
  device = ids[0];
}

// ---------------------------------------------------------------------------
// materials, cross sections, attenuations
// ---------------------------------------------------------------------------
void GGEMSMaterials::add(const std::string& name) {
  if (names.size() == 255) raise("GGEMSMaterials", "Limit of material reached. The limit is 255 materials!!!");
  names.push_back(name);
}

void GGEMSMaterials::set_distance_cut(const std::string& particle, float value, const std::string& unit) {
  float cut = distance_unit(value, unit);
  if (particle == "gamma") {
    photon_distance_cut = cut;
    if (cut < 1.e-3f) raise("GGEMSRangeCuts", "Distance cut for photon is too small!!! Minimum value is 0.001 mm!!!");
  } else if (particle == "e-" || particle == "e+") {
    if (cut < 1.e-3f) raise("GGEMSRangeCuts", "Distance cut is too small!!! Minimum value is 0.001 mm!!!");
    // electron/positron cuts have no effect on photon-only transport
  } else {
    raise("GGEMSMaterials", "Particle name " + particle + " unknown!!! The particles are: gamma, e-, e+");
  }
}

void GGEMSMaterials::initialize() {
  tables = build_material_tables(names);
  const float emin = GGEMSProcessesManager::get().emin;
  for (size_t i = 0; i < names.size(); ++i)
    tables.photon_cut[i] = photon_energy_cut(tables, (int)i, photon_distance_cut, emin);
  built = true;
}

int GGEMSMaterials::index(const std::string& name) const {
  for (size_t i = 0; i < names.size(); ++i)
    if (names[i] == name) return (int)i;
  raise("GGEMSMaterials", "Material '" + name + "' not found in the navigator materials!!!");
}

void GGEMSCrossSections::add_process(const std::string& process, const std::string& particle) {
  if (particle != "gamma") raise("GGEMSCrossSections", "Only 'gamma' processes are transported");
  int id = process_id(process);
  if (std::find(processes.begin(), processes.end(), id) != processes.end()) {
    log(3, "GGEMSCrossSections", process + " process already activated!!!");
    return;
  }
  processes.push_back(id);
}

void GGEMSCrossSections::initialize() {
  const GGEMSProcessesManager& pm = GGEMSProcessesManager::get();
  if (!materials->built) materials->initialize();
  tables = build_cross_sections(materials->tables, processes, pm.nbins, pm.emin, pm.emax);
  built = true;
}

void GGEMSAttenuations::initialize() {
  if (!materials->built) materials->initialize();
  tables = build_attenuations(materials->tables);
  built = true;
}

// ---------------------------------------------------------------------------
// navigators
// ---------------------------------------------------------------------------
GGEMSNavigator::GGEMSNavigator(const std::string& n) : name(n) {
  GGEMSNavigatorManager::get().navigators.push_back(this);
}

GGEMSNavigator::~GGEMSNavigator() {
  auto& v = GGEMSNavigatorManager::get().navigators;
  v.erase(std::remove(v.begin(), v.end(), this), v.end());
}

void GGEMSNavigator::set_position(float x, float y, float z, const std::string& unit) {
  update_pos = true;
  position[0] = distance_unit(x, unit);
  position[1] = distance_unit(y, unit);
  position[2] = distance_unit(z, unit);
}

void GGEMSNavigator::set_rotation(float x, float y, float z, const std::string& unit) {
  update_rot = true;
  rotation[0] = angle_unit(x, unit);
  rotation[1] = angle_unit(y, unit);
  rotation[2] = angle_unit(z, unit);
}

namespace {
// GGEMSVoxelizedSolid::ConvertImageToLabel (include/GGEMS/geometries/GGEMSVoxelizedSolid.hh:161-257)
template <typename T>
void label_image(const std::string& raw, size_t n, const std::string& range_file, GGEMSMaterials& mats,
                 std::vector<uint8_t>& labels) {
  std::ifstream in(raw, std::ios::binary);
  if (!in) raise("GGEMSVoxelizedSolid", "Problem reading filename: " + raw);
  std::vector<T> v(n);
  in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)));
  if ((size_t)in.gcount() != n * sizeof(T)) raise("GGEMSVoxelizedSolid", "raw file too short: " + raw);
  labels.assign(n, 0xFF);
  std::ifstream rf(range_file);
  if (!rf) raise("GGEMSVoxelizedSolid", "Problem reading filename: " + range_file);
  std::string line, material;
  int label = 0;
  while (std::getline(rf, line)) {
    if (line.find_first_not_of("\t\n ") == std::string::npos) continue;
    std::istringstream iss(line);
    float first = 0.0f, last = 0.0f;
    iss >> first >> last >> material;
    mats.add(material);
    for (size_t i = 0; i < n; ++i) {
      float value = static_cast<float>(v[i]);
      if ((value == first && value == last) || (value >= first && value < last)) labels[i] = (uint8_t)label;
    }
    ++label;
  }
  for (uint8_t l : labels)
    if (l == 0xFF) raise("GGEMSVoxelizedSolid", "Errors(s) in the range data file!!!");
}
}  // namespace

void GGEMSVoxelizedPhantom::initialize() {
  if (phantom_file.empty() || range_file.empty())
    raise("GGEMSVoxelizedPhantom", "You have to set a mhd file and a range data file!!!");
  MHDHeader h = read_mhd_header(phantom_file);
  for (int i = 0; i < 3; ++i) {
    dims[i] = h.dim[i];
    voxel_size[i] = h.spacing[i];
    // GGEMSMHDImage::Read borders (src/io/GGEMSMHDImage.cc:206-209)
    border_min[i] = -static_cast<float>(dims[i]) * voxel_size[i] * 0.5f;
    border_max[i] = static_cast<float>(dims[i]) * voxel_size[i] * 0.5f;
  }
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  materials.names.clear();
  if (h.type == "MET_CHAR") label_image<int8_t>(h.raw_file, n, range_file, materials, labels);
  else if (h.type == "MET_UCHAR") label_image<uint8_t>(h.raw_file, n, range_file, materials, labels);
  else if (h.type == "MET_SHORT") label_image<int16_t>(h.raw_file, n, range_file, materials, labels);
  else if (h.type == "MET_USHORT") label_image<uint16_t>(h.raw_file, n, range_file, materials, labels);
  else if (h.type == "MET_INT") label_image<int32_t>(h.raw_file, n, range_file, materials, labels);
  else if (h.type == "MET_UINT") label_image<uint32_t>(h.raw_file, n, range_file, materials, labels);
  else if (h.type == "MET_FLOAT") label_image<float>(h.raw_file, n, range_file, materials, labels);
  else if (h.type == "MET_DOUBLE") label_image<double>(h.raw_file, n, range_file, materials, labels);
  else raise("GGEMSMHDImage", "Value invalid for the key 'ElementType'!!!");
  // rotation before position (src/navigators/GGEMSVoxelizedPhantom.cc:120-135)
  static const float axis[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  transform = Transform();
  transform.set_axis(axis);
  if (update_rot) transform.rotate(rotation[0], rotation[1], rotation[2]);
  if (update_pos) transform.translate(position[0], position[1], position[2]);
  materials.initialize();
  cross_sections.initialize();
}

int GGEMSVoxelizedPhantom::register_solids(ggh_context* ctx, int material_offset) {
  ggh_voxel_solid_desc d;
  std::memset(&d, 0, sizeof d);
  std::memcpy(d.obb.matrix, transform.m.data(), sizeof(float) * 16);
  for (int i = 0; i < 3; ++i) {
    d.obb.border_min[i] = border_min[i];
    d.obb.border_max[i] = border_max[i];
    d.n_voxels[i] = dims[i];
    d.voxel_size[i] = voxel_size[i];
  }
  d.material_offset = material_offset;
  d.labels = labels.data();
  solid_index = ggh_add_voxel_solid(ctx, &d);
  engine_check(solid_index, "register voxelized phantom");
  return 1;
}

void GGEMSVoxelizedPhantom::save_results(ggh_context* ctx) {
  if (dosimetry) dosimetry->save(ctx);
}

bool GGEMSVoxelizedPhantom::dose_geometry(int& solid, float vs[3], float lo[3], float hi[3]) const {
  solid = solid_index;
  for (int i = 0; i < 3; ++i) {
    vs[i] = voxel_size[i];
    lo[i] = border_min[i];
    hi[i] = border_max[i];
  }
  return true;
}

// ---- GGEMSMeshedPhantom -------------------------------------------------------------------------
namespace {

// GGEMSSTLReader::Read (src/io/GGEMSSTLReader.cc:71-99): binary STL, the
// three vertices of each record (the normal and attribute are skipped)
std::vector<float> read_stl(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) raise("GGEMSSTLReader", "Problem reading filename: " + path);
  char header[80];
  uint32_t n = 0;
  f.read(header, 80);
  f.read(reinterpret_cast<char*>(&n), 4);
  if (!f || n == 0) raise("GGEMSSTLReader", "Invalid STL file: " + path);
  std::vector<float> tri((size_t)n * 9);
  for (uint32_t i = 0; i < n; ++i) {
    float data[12];
    uint16_t attr;
    f.read(reinterpret_cast<char*>(data), sizeof data);
    f.read(reinterpret_cast<char*>(&attr), sizeof attr);
    if (!f) raise("GGEMSSTLReader", "Truncated STL file: " + path);
    for (int k = 0; k < 9; ++k) tri[(size_t)9 * i + k] = data[3 + k];
  }
  return tri;
}

}  // namespace

void GGEMSMeshedPhantom::initialize() {
  // GGEMSMeshedPhantom::CheckParameters / Initialize (src/navigators/GGEMSMeshedPhantom.cc:64-155)
  if (phantom_file.empty()) raise("GGEMSMeshedPhantom", "You have to set a mesh file in STL format!!!");
  if (octree_depth > 8) raise("GGEMSOctree", "Max. depth too large in Octree!!!");
  if (octree_depth < 1) raise("GGEMSOctree", "Octree depth has to be at least 1!!!");
  if (materials.names.empty()) raise("GGEMSMeshedPhantom", "A material has to be set for the meshed phantom!!!");
  triangles = read_stl(phantom_file);
  // rotation before position; UpdateTriangles applies the rows of the
  // matrix (dot product, left to right) plus its translation to every point
  // (src/geometries/GGEMSMeshedSolid.cc:205-262)
  static const float axis[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  Transform t;
  t.set_axis(axis);
  if (update_rot) t.rotate(rotation[0], rotation[1], rotation[2]);
  if (update_pos) t.translate(position[0], position[1], position[2]);
  const float* m = t.m.data();
  for (size_t k = 0; k < triangles.size(); k += 3) {
    const float x = triangles[k], y = triangles[k + 1], z = triangles[k + 2];
    for (int r = 0; r < 3; ++r) {
      float v = m[4 * r] * x;
      v += m[4 * r + 1] * y;
      v += m[4 * r + 2] * z;
      triangles[k + r] = v + m[4 * r + 3];
    }
  }
  // ComputeOctreeCenter / ComputeHalfWidthCenter (src/geometries/GGEMSMeshedSolid.cc:340-458):
  // bounds (the maximum starts at FLT_MIN), grown by 2.5 %, rounded outwards to whole mm
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {FLT_MIN, FLT_MIN, FLT_MIN};
  for (size_t k = 0; k < triangles.size(); k += 3)
    for (int a = 0; a < 3; ++a) {
      lo[a] = triangles[k + a] < lo[a] ? triangles[k + a] : lo[a];
      hi[a] = triangles[k + a] > hi[a] ? triangles[k + a] : hi[a];
    }
  for (int a = 0; a < 3; ++a) {
    const float e = (hi[a] - lo[a]) * 0.025f;
    lo[a] -= e;
    hi[a] += e;
    lo[a] = lo[a] < 0.0f ? std::floor(lo[a]) : std::ceil(lo[a]);
    hi[a] = hi[a] < 0.0f ? std::floor(hi[a]) : std::ceil(hi[a]);
    const float center = (hi[a] + lo[a]) * 0.5f, half = (hi[a] - lo[a]) * 0.5f;
    obb_min[a] = center - half;
    obb_max[a] = center + half;
  }
  materials.initialize();
  cross_sections.initialize();
}

int GGEMSMeshedPhantom::register_solids(ggh_context* ctx, int material_offset) {
  ggh_mesh_solid_desc d;
  std::memset(&d, 0, sizeof d);
  d.obb.matrix[0] = d.obb.matrix[5] = d.obb.matrix[10] = d.obb.matrix[15] = 1.0f;
  for (int i = 0; i < 3; ++i) {
    d.obb.border_min[i] = obb_min[i];
    d.obb.border_max[i] = obb_max[i];
  }
  d.material_offset = material_offset;
  d.n_triangles = (int32_t)(triangles.size() / 9);
  d.triangles = triangles.data();
  solid_index = ggh_add_mesh_solid(ctx, &d);
  engine_check(solid_index, "register meshed phantom");
  return 1;
}

void GGEMSMeshedPhantom::save_results(ggh_context* ctx) {
  if (dosimetry) dosimetry->save(ctx);
}

// GGEMSMeshedSolid::GetVoxelSizes is zero; the dose map spans the octree box
bool GGEMSMeshedPhantom::dose_geometry(int& solid, float vs[3], float lo[3], float hi[3]) const {
  solid = solid_index;
  for (int i = 0; i < 3; ++i) {
    vs[i] = 0.0f;
    lo[i] = obb_min[i];
    hi[i] = obb_max[i];
  }
  return true;
}

void GGEMSCTSystem::initialize() {
  if (type != "curved" && type != "flat") raise("GGEMSCTSystem", "Available CT system types: 'curved' or 'flat'");
  if (sdd == 0.0f) raise("GGEMSCTSystem", "For CT system, source detector distance (SDD) has to be > 0.0 mm!!!");
  if (sid > sdd) raise("GGEMSCTSystem", "Source isocenter distance must be inferior to source detector distance!!!");
  if (modules[0] == 0 || modules[1] == 0)
    raise("GGEMSSystem", "In system parameters, number of module in x and y axis (local axis) has to be > 0!!!");
  if (elements[0] == 0 || elements[1] == 0 || elements[2] == 0)
    raise("GGEMSSystem", "In system parameters, number of detection elements in x, y and z axis (local axis) has to be > 0!!!");
  if (element_size[0] == 0.0f || element_size[1] == 0.0f || element_size[2] == 0.0f)
    raise("GGEMSSystem", "In system parameters, size of detection elements (local axis) has to be > 0.0 mm!!!");
  if (materials.names.empty()) raise("GGEMSSystem", "In system parameters, a material has to be defined!!!");
  if (output_basename.empty()) raise("GGEMSNavigator", "Output basename not set!!!");
  // GGEMSSolidBox local axis (src/geometries/GGEMSSolidBox.cc:71-77)
  static const float axis[9] = {0, 0, 1, 0, 1, 0, -1, 0, 0};
  const size_t nmod = modules[0] * modules[1];
  module_transforms.assign(nmod, Transform());
  for (auto& t : module_transforms) t.set_axis(axis);
  if (type == "curved") {
    // InitializeCurvedGeometry (src/navigators/GGEMSCTSystem.cc:131-183)
    float c = (static_cast<float>(elements[1]) * element_size[1]) * 0.5f;
    float rho = std::sqrt(sdd * sdd + c * c);
    float alpha = 2.0f * std::asin(c / rho);
    float ox = -sid, oy = 0.0f, px = sdd - sid, py = 0.0f;
    for (size_t j = 0; j < modules[1]; ++j) {
      float step = alpha * (static_cast<float>(j) + 0.5f * (1.0f - static_cast<float>(modules[1])));
      float gx = (px - ox) * std::cos(step) - (py - oy) * std::sin(step) + ox;
      float gy = (px - ox) * std::sin(step) + (py - oy) * std::cos(step) + oy;
      for (size_t i = 0; i < modules[0]; ++i) {
        float gz = static_cast<float>(elements[0]) * element_size[0] *
                   (static_cast<float>(i) + 0.5f * (1.0f - static_cast<float>(modules[0])));
        Transform& t = module_transforms[i + j * modules[0]];
        t.rotate(0.0f, 0.0f, step);
        t.translate(gx + global_position[0], gy + global_position[1], gz + global_position[2]);
      }
    }
  } else {
    // InitializeFlatGeometry (src/navigators/GGEMSCTSystem.cc:186-222)
    float gx = sdd - sid;
    for (size_t j = 0; j < modules[1]; ++j) {
      float gy = static_cast<float>(elements[1]) * element_size[1] *
                 (static_cast<float>(j) + 0.5f * (1.0f - static_cast<float>(modules[1])));
      for (size_t i = 0; i < modules[0]; ++i) {
        float gz = static_cast<float>(elements[0]) * element_size[0] *
                   (static_cast<float>(i) + 0.5f * (1.0f - static_cast<float>(modules[0])));
        Transform& t = module_transforms[i + j * modules[0]];
        t.rotate(0.0f, 0.0f, 0.0f);
        t.translate(gx + global_position[0], gy + global_position[1], gz + global_position[2]);
      }
    }
  }
  if (update_rot)
    for (auto& t : module_transforms) t.rotate(rotation[0], rotation[1], rotation[2]);
  materials.initialize();
  cross_sections.initialize();
}

int GGEMSCTSystem::register_solids(ggh_context* ctx, int material_offset) {
  solid_indices.clear();
  const float size[3] = {static_cast<float>(elements[0]) * element_size[0],
                         static_cast<float>(elements[1]) * element_size[1],
                         static_cast<float>(elements[2]) * element_size[2]};
  for (const Transform& t : module_transforms) {
    ggh_box_solid_desc d;
    std::memset(&d, 0, sizeof d);
    std::memcpy(d.obb.matrix, t.m.data(), sizeof(float) * 16);
    for (int i = 0; i < 3; ++i) {
      d.obb.border_min[i] = -size[i] * 0.5f;
      d.obb.border_max[i] = size[i] * 0.5f;
      d.n_elements[i] = (int32_t)elements[i];
      d.box_size[i] = size[i];
    }
    d.material_offset = material_offset;
    d.threshold = threshold;
    d.store_scatter = store_scatter ? 1 : 0;
    int k = ggh_add_box_solid(ctx, &d);
    engine_check(k, "register CT module");
    solid_indices.push_back(k);
  }
  return (int)module_transforms.size();
}

// GGEMSSystem::SaveResults (src/navigators/GGEMSSystem.cc:165-256)
void GGEMSCTSystem::save_results(ggh_context* ctx) {
  const size_t ex = elements[0], ey = elements[1], ez = elements[2];
  const int dim[3] = {(int)(modules[0] * ex), (int)(modules[1] * ey), (int)ez};
  const size_t total = (size_t)dim[0] * dim[1] * dim[2];
  std::vector<int32_t> out(total, 0), sout(store_scatter ? total : 0, 0), h(ex * ey), s(ex * ey);
  for (size_t jj = 0; jj < modules[1]; ++jj)
    for (size_t ii = 0; ii < modules[0]; ++ii) {
      int k = solid_indices[ii + jj * modules[0]];
      engine_check(ggh_get_histogram(ctx, k, h.data(), store_scatter ? s.data() : nullptr), "read histogram");
      for (size_t jjj = 0; jjj < ey; ++jjj)
        for (size_t iii = 0; iii < ex; ++iii) {
          size_t o = (iii + ii * ex) + (jjj + jj * ey) * (size_t)dim[0];
          out[o] += h[iii + jjj * ex];
          if (store_scatter) sout[o] += s[iii + jjj * ex];
        }
    }
  write_mhd(output_basename, "MET_INT", dim, element_size, out.data(), sizeof(int32_t));
  if (store_scatter) {
    std::string sname = output_basename;
    size_t f = sname.find(".mhd");
    sname = f == std::string::npos ? sname + "-scatter.mhd" : sname.substr(0, f) + "-scatter.mhd";
    write_mhd(sname, "MET_INT", dim, element_size, sout.data(), sizeof(int32_t));
  }
}

// the projections of an acquisition, one MHD stack per CT system (same
// module layout as SaveResults, one image per projection)
void GGEMSCTSystem::save_acquisition(ggh_context* ctx, int n_proj) {
  const size_t ex = elements[0], ey = elements[1];
  const int dim[3] = {(int)(modules[0] * ex), (int)(modules[1] * ey), n_proj};
  const size_t plane = (size_t)dim[0] * dim[1];
  std::vector<int32_t> out(plane * n_proj, 0), sout(store_scatter ? plane * n_proj : 0, 0), h(ex * ey), s(ex * ey);
  for (int p = 0; p < n_proj; ++p)
    for (size_t jj = 0; jj < modules[1]; ++jj)
      for (size_t ii = 0; ii < modules[0]; ++ii) {
        int k = solid_indices[ii + jj * modules[0]];
        engine_check(ggh_get_projection(ctx, k, p, h.data(), store_scatter ? s.data() : nullptr), "read projection");
        for (size_t jjj = 0; jjj < ey; ++jjj)
          for (size_t iii = 0; iii < ex; ++iii) {
            size_t o = (size_t)p * plane + (iii + ii * ex) + (jjj + jj * ey) * (size_t)dim[0];
            out[o] = h[iii + jjj * ex];
            if (store_scatter) sout[o] = s[iii + jjj * ex];
          }
      }
  write_mhd(output_basename, "MET_INT", dim, element_size, out.data(), sizeof(int32_t));
  if (store_scatter) {
    std::string sname = output_basename;
    size_t f = sname.find(".mhd");
    sname = f == std::string::npos ? sname + "-scatter.mhd" : sname.substr(0, f) + "-scatter.mhd";
    write_mhd(sname, "MET_INT", dim, element_size, sout.data(), sizeof(int32_t));
  }
}

// ---------------------------------------------------------------------------
// dosimetry (src/navigators/GGEMSDosimetryCalculator.cc:435-809)
// ---------------------------------------------------------------------------
void GGEMSDosimetryCalculator::initialize(ggh_context* ctx) {
  if (!navigator) raise("GGEMSDosimetryCalculator", "A navigator has to be associated to GGEMSDosimetryCalculator!!!");
  int solid = -1;
  float nav_vs[3], lo[3], hi[3];
  if (!navigator->dose_geometry(solid, nav_vs, lo, hi))
    raise("GGEMSDosimetryCalculator", "dosimetry attaches to a voxelized or meshed phantom");
  float vs[3] = {dosel_size[0], dosel_size[1], dosel_size[2]};
  if (dosel_size[0] < 0.0f && dosel_size[1] < 0.0f && dosel_size[2] < 0.0f)
    for (int i = 0; i < 3; ++i) vs[i] = nav_vs[i];
  if (photon_tracking)
    for (int i = 0; i < 3; ++i)
      if (vs[i] != nav_vs[i])
        raise("GGEMSDosimetryCalculator",
              "Dosel size and voxel size in voxelized phantom have to be the same when photon tracking is activated!!!");
  // a meshed solid has no voxel size: without a dosel size its dose map
  // would have a zero-sized dosel
  for (int i = 0; i < 3; ++i)
    if (!(vs[i] > 0.0f)) raise("GGEMSDosimetryCalculator", "Dosel size has to be set for a meshed phantom!!!");
  ggh_dosimetry_desc d;
  std::memset(&d, 0, sizeof d);
  d.solid = solid;
  for (int i = 0; i < 3; ++i) {
    used_size[i] = vs[i];
    d.dosel_size[i] = vs[i];
    float extent = hi[i] - lo[i];
    n_dosels[i] = (int)static_cast<size_t>(extent / vs[i]);
    d.n_dosels[i] = n_dosels[i];
    d.border_min[i] = lo[i];
    d.border_max[i] = hi[i];
  }
  d.edep_squared = (edep_squared || uncertainty) ? 1 : 0;
  d.hit = (hit || uncertainty) ? 1 : 0;
  d.photon_tracking = photon_tracking ? 1 : 0;
  engine_check(ggh_set_dosimetry(ctx, &d), "set dosimetry");
}

void GGEMSDosimetryCalculator::save(ggh_context* ctx) {
  const size_t n = (size_t)n_dosels[0] * n_dosels[1] * n_dosels[2];
  const bool want_e2 = edep_squared || uncertainty, want_hit = hit || uncertainty;
  std::vector<double> e(n), e2(want_e2 ? n : 0);
  std::vector<int32_t> hits(want_hit ? n : 0), pt(photon_tracking ? n : 0);
  engine_check(ggh_get_dose_tallies(ctx, e.data(), want_e2 ? e2.data() : nullptr, want_hit ? hits.data() : nullptr,
                                    photon_tracking ? pt.data() : nullptr),
               "read dose tallies");
  std::vector<float> dose(n), unc(uncertainty ? n : 0);
  engine_check(ggh_compute_dose(ctx, scale, water_reference ? 1 : 0, minimum_density, dose.data(),
                                uncertainty ? unc.data() : nullptr),
               "compute dose");
  write_mhd(output + "_dose.mhd", "MET_FLOAT", n_dosels, used_size, dose.data(), sizeof(float));
  if (photon_tracking) write_mhd(output + "_photon_tracking.mhd", "MET_INT", n_dosels, used_size, pt.data(), 4);
  if (edep) write_mhd(output + "_edep.mhd", "MET_DOUBLE", n_dosels, used_size, e.data(), 8);
  if (hit) write_mhd(output + "_hit.mhd", "MET_INT", n_dosels, used_size, hits.data(), 4);
  if (edep_squared) write_mhd(output + "_edep_squared.mhd", "MET_DOUBLE", n_dosels, used_size, e2.data(), 8);
  if (uncertainty) write_mhd(output + "_uncertainty.mhd", "MET_FLOAT", n_dosels, used_size, unc.data(), 4);
}

// ---------------------------------------------------------------------------
// X-ray source
// ---------------------------------------------------------------------------
GGEMSXRaySource::GGEMSXRaySource(const std::string& n) : name(n) {
  // local axis of the X-ray source (src/sources/GGEMSXRaySource.cc:45-51)
  static const float axis[9] = {0, 0, -1, 0, 1, 0, 1, 0, 0};
  transform.set_axis(axis);
  GGEMSSourceManager::get().sources.push_back(this);
}

GGEMSXRaySource::~GGEMSXRaySource() {
  auto& v = GGEMSSourceManager::get().sources;
  v.erase(std::remove(v.begin(), v.end(), this), v.end());
}

void GGEMSXRaySource::set_position(float x, float y, float z, const std::string& unit) {
  transform.translate(distance_unit(x, unit), distance_unit(y, unit), distance_unit(z, unit));
  position_set = true;
}

void GGEMSXRaySource::set_rotation(float x, float y, float z, const std::string& unit) {
  transform.rotate(angle_unit(x, unit), angle_unit(y, unit), angle_unit(z, unit));
  rotation_set = true;
}

void GGEMSXRaySource::set_monoenergy(float e, const std::string& unit) {
  spectrum.clear();
  spectrum.push_back({energy_unit(e, unit), 1.0});
  is_interp = false;
}

void GGEMSXRaySource::set_polyenergy(const std::string& file) {
  spectrum.clear();
  std::ifstream in(file);
  if (!in) raise("GGEMSSource", "Problem reading filename: " + file);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    double e = 0.0, w = 0.0;
    iss >> e >> w;
    spectrum.push_back({e, w});
  }
}

void GGEMSXRaySource::set_energy_peak(float e, const std::string& unit, float intensity) {
  spectrum.push_back({energy_unit(e, unit), (double)intensity});
  is_interp = false;
}

void GGEMSXRaySource::check() const {
  if (particle_type == 99) raise("GGEMSSource", "You have to set a particle type for the source: gamma");
  if (particle_type != 0) raise("GGEMSSource", "Only 'gamma' sources are transported");
  if (name.empty()) raise("GGEMSSource", "You have to set a name for the source!!!");
  if (n_particles == 0) raise("GGEMSSource", "You have to set a number of particles > 0!!!");
  if (!position_set) raise("GGEMSSource", "You have to set a position for the source!!!");
  if (!rotation_set) raise("GGEMSSource", "You have to set a rotation for the source!!!");
  if (!aperture_set) raise("GGEMSXRaySource", "You have to set a beam aperture for the source!!!");
  if (aperture < 0.0f) raise("GGEMSXRaySource", "The beam aperture must be >= 0!!!");
  if (!focal_set) raise("GGEMSXRaySource", "You have to set a focal spot size!!!");
  if (focal[0] < 0.0f || focal[1] < 0.0f || focal[2] < 0.0f)
    raise("GGEMSXRaySource", "The focal spot size is a posivite value!!!");
  if (spectrum.empty()) raise("GGEMSSource", "You have to set an energy (mono, poly or peaks)!!!");
}

void GGEMSXRaySource::fill_energy(std::vector<float>& energy, std::vector<float>& cdf) const {
  const size_t nb = spectrum.size();
  energy.assign(nb + 1, 0.0f);
  cdf.assign(nb + 1, 0.0f);
  double sum = 0.0;
  for (auto& s : spectrum) sum += s.second;
  std::vector<double> tmp(nb + 1, 0.0);
  for (size_t i = 0; i < nb; ++i) {
    tmp[i + 1] = spectrum[i].second;
    energy[i] = static_cast<float>(spectrum[i].first);
  }
  energy[nb] = energy[nb - 1];
  cdf[0] = static_cast<float>(tmp[0]);
  for (size_t i = 1; i <= nb; ++i) {
    tmp[i] = tmp[i] + tmp[i - 1];
    cdf[i] = static_cast<float>(tmp[i] / sum);
  }
  cdf[nb] = 1.0f;
}

// ---------------------------------------------------------------------------
// GGEMS
// ---------------------------------------------------------------------------
GGEMS::~GGEMS() {
  if (ctx) {
    if (g_current_ctx == ctx) g_current_ctx = nullptr;
    ggh_context_destroy(ctx);
    ctx = nullptr;
  }
}

// GGEMS::Initialize (src/global/GGEMS.cc:199-297)
void GGEMS::initialize(uint32_t s) {
  if (!MaterialDatabase::get().ready()) raise("GGEMS", "Materials are not loaded in GGEMS!!!");
  auto& sources = GGEMSSourceManager::get().sources;
  auto& navs = GGEMSNavigatorManager::get().navigators;
  if (sources.empty()) raise("GGEMS", "You have to define a source!!!");
  for (GGEMSXRaySource* src : sources) src->check();
  if (navs.empty()) log(0, "GGEMSNavigatorManager", "You have not defined a navigator (detector or phantom).");
  if (ctx) {
    if (g_current_ctx == ctx) g_current_ctx = nullptr;
    ggh_context_destroy(ctx);
    ctx = nullptr;
  }
  engine_check(ggh_context_create(GGEMSOpenCLManager::get().device, &ctx), "create engine context");
  g_current_ctx = ctx;
  // navigators: geometry, materials, cuts, cross sections
  for (GGEMSNavigator* n : navs) n->initialize();
  // all navigators share one process list: the engine draws one uniform per listed process
  std::vector<int> procs = navs.empty() ? std::vector<int>{} : navs[0]->cross_sections.processes;
  for (GGEMSNavigator* n : navs)
    if (n->cross_sections.processes != procs)
      raise("GGEMS", "every navigator must activate the same photon processes in the same order");
  // concatenate the navigators' tables into the engine's global material
  // space.  A navigator whose tables equal an earlier navigator's in every
  // value (same materials in the same order, same cuts; e.g. several
  // phantoms of one acquisition) shares that navigator's materials, which
  // keeps the step tables small enough to stage in LDS.
  const GGEMSProcessesManager& pm = GGEMSProcessesManager::get();
  const int nb = pm.nbins;
  auto same_tables = [](const GGEMSNavigator* a, const GGEMSNavigator* b) {
    const MaterialTables &x = a->materials.tables, &y = b->materials.tables;
    return x.names == y.names && x.n_elements == y.n_elements && x.density == y.density &&
           x.photon_cut == y.photon_cut && x.z == y.z && x.atom_density == y.atom_density &&
           a->cross_sections.tables.cs == b->cross_sections.tables.cs;
  };
  std::vector<int> shares(navs.size(), -1);  // the earlier navigator whose materials a navigator reuses
  int nmat = 0;
  for (size_t i = 0; i < navs.size(); ++i) {
    for (size_t j = 0; j < i && shares[i] < 0; ++j)
      if (shares[j] < 0 && same_tables(navs[i], navs[j])) shares[i] = (int)j;
    if (shares[i] < 0) nmat += (int)navs[i]->materials.names.size();
  }
  if (nmat == 0) raise("GGEMS", "no material defined in the navigators");
  std::vector<float> cs((size_t)3 * nmat * nb, 0.0f), cs_atom((size_t)101 * nb, 0.0f), ad, cut, rho;
  std::vector<int32_t> n_el, el_idx;
  std::vector<uint8_t> z;
  std::vector<int> offsets;
  int moff = 0;
  for (size_t i = 0; i < navs.size(); ++i) {
    GGEMSNavigator* n = navs[i];
    if (shares[i] >= 0) {
      offsets.push_back(offsets[shares[i]]);
      continue;
    }
    offsets.push_back(moff);
    const MaterialTables& mt = n->materials.tables;
    const CrossSectionTables& ct = n->cross_sections.tables;
    const int nm = (int)mt.names.size();
    for (int m = 0; m < nm; ++m) {
      n_el.push_back(mt.n_elements[m]);
      el_idx.push_back((int32_t)z.size() + mt.element_index[m] - mt.element_index[0]);
      cut.push_back(mt.photon_cut[m]);
      rho.push_back(mt.density[m]);
      for (int p = 0; p < 3; ++p)
        std::memcpy(&cs[((size_t)p * nmat + moff + m) * nb], &ct.cs[((size_t)p * nm + m) * nb], sizeof(float) * nb);
    }
    for (int k = 0; k < mt.total_elements(); ++k) {
      z.push_back(mt.z[k]);
      ad.push_back(mt.atom_density[k]);
      std::memcpy(&cs_atom[(size_t)mt.z[k] * nb], &ct.cs_atom[((size_t)kRayleigh * 101 + mt.z[k]) * nb],
                  sizeof(float) * nb);
    }
    moff += nm;
  }
  const PhysicsData& pd = PhysicsData::get();
  std::vector<float> pp((size_t)9 * 101);
  for (int k = 0; k < 9; ++k) {
    size_t cnt = 0;
    const float* a = pd.f32("rayleigh_pp" + std::to_string(k), &cnt);
    std::memcpy(&pp[(size_t)k * 101], a, sizeof(float) * 101);
  }
  std::vector<float> ebins = energy_grid(nb, pm.emin, pm.emax);
  ggh_physics_desc d;
  std::memset(&d, 0, sizeof d);
  d.nbins = nb;
  d.n_materials = nmat;
  d.n_processes = (int32_t)procs.size();
  for (size_t i = 0; i < procs.size(); ++i) d.process_id[i] = procs[i];
  d.energy_bins = ebins.data();
  d.cross_sections = cs.data();
  d.cs_atom_rayleigh = cs_atom.data();
  d.total_elements = (int32_t)z.size();
  d.n_elements = n_el.data();
  d.element_index = el_idx.data();
  d.atomic_number = z.data();
  d.atom_density = ad.data();
  d.photon_cut = cut.data();
  d.density = rho.data();
  d.rayleigh_pp = pp.data();
  d.rayleigh_factor = *pd.f32("rayleigh_factor");
  engine_check(ggh_set_physics(ctx, &d), "upload physics tables");
  for (size_t i = 0; i < navs.size(); ++i) navs[i]->register_solids(ctx, offsets[i]);
  for (GGEMSNavigator* n : navs)
    if (n->dosimetry) n->dosimetry->initialize(ctx);
  seed = s;
  engine_check(ggh_seed(ctx, s), "seed particle slots");
  upload_source(*sources.front());
  initialized = true;
  log(0, "GGEMS", "GGEMS initialization succeeded");
}

void GGEMS::upload_source(const GGEMSXRaySource& src) {
  std::vector<float> energy, cdf;
  src.fill_energy(energy, cdf);
  ggh_xray_source_desc d;
  std::memset(&d, 0, sizeof d);
  std::memcpy(d.matrix, src.transform.m.data(), sizeof(float) * 16);
  d.aperture = src.aperture;
  for (int i = 0; i < 3; ++i) d.focal_spot[i] = src.focal[i];
  d.n_energy = (int32_t)src.spectrum.size();
  d.energy = energy.data();
  d.cdf = cdf.data();
  d.is_interp = src.is_interp ? 1 : 0;
  engine_check(ggh_set_xray_source(ctx, &d), "set source");
}

// GGEMS::Run (src/global/GGEMS.cc:303-414): every source's photons, then dose and outputs
void GGEMS::run() {
  if (!initialized) raise("GGEMS", "GGEMS is not initialized!!!");
  auto t0 = std::chrono::steady_clock::now();
  for (GGEMSXRaySource* src : GGEMSSourceManager::get().sources) {
    upload_source(*src);
    engine_check(ggh_run(ctx, src->n_particles), "transport");
  }
  auto t1 = std::chrono::steady_clock::now();
  last_run_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  for (GGEMSNavigator* n : GGEMSNavigatorManager::get().navigators) n->save_results(ctx);
  log(0, "GGEMS", "GGEMS simulation succeeded");
}

void GGEMS::run_acquisition(size_t n, const float* rx, const float* ry, const float* rz, const uint64_t* np,
                            const int32_t* phantom) {
  if (!initialized) raise("GGEMS", "GGEMS is not initialized!!!");
  auto& sources = GGEMSSourceManager::get().sources;
  if (sources.size() != 1) raise("GGEMS", "A CT acquisition needs exactly one X-ray source!!!");
  if (n == 0 || n > GGH_MAX_PROJECTIONS) raise("GGEMS", "Invalid number of projections!!!");
  // module placements in engine registration order (CT systems in navigator order)
  std::vector<const Transform*> modules;
  for (GGEMSNavigator* nav : GGEMSNavigatorManager::get().navigators)
    if (auto* ct = dynamic_cast<GGEMSCTSystem*>(nav))
      for (const Transform& t : ct->module_transforms) modules.push_back(&t);
  std::vector<float> src_m(16 * n), box_m(16 * n * modules.size());
  for (size_t p = 0; p < n; ++p) {
    Transform t = sources.front()->transform;
    t.rotate(rx[p], ry[p], rz[p]);
    std::memcpy(&src_m[16 * p], t.m.data(), sizeof(float) * 16);
    for (size_t j = 0; j < modules.size(); ++j) {
      Transform u = *modules[j];
      u.rotate(rx[p], ry[p], rz[p]);
      std::memcpy(&box_m[16 * (p * modules.size() + j)], u.m.data(), sizeof(float) * 16);
    }
  }
  // phantom navigator ranks -> engine solids
  std::vector<int32_t> only;
  if (phantom) {
    std::vector<int> solids;
    for (GGEMSNavigator* nav : GGEMSNavigatorManager::get().navigators)
      if (nav->phantom_solid() >= 0) solids.push_back(nav->phantom_solid());
    only.resize(n);
    for (size_t p = 0; p < n; ++p) {
      if (phantom[p] < -1 || phantom[p] >= (int32_t)solids.size())
        raise("GGEMS", "Projection " + std::to_string(p) + " names no phantom navigator!!!");
      only[p] = phantom[p] < 0 ? -1 : solids[phantom[p]];
    }
  }
  upload_source(*sources.front());
  ggh_acquisition_desc d;
  d.n_projections = (int32_t)n;
  d.source_matrices = src_m.data();
  d.box_matrices = box_m.data();
  d.n_particles = np;
  d.phantoms = only.empty() ? nullptr : only.data();
  engine_check(ggh_set_acquisition(ctx, &d), "set acquisition");
  acquisition_projections = (int)n;
  engine_check(ggh_capture_acquisition(ctx), "capture acquisition");
  repeat_acquisition();
}

void GGEMS::repeat_acquisition() {
  if (!acquisition_projections) raise("GGEMS", "No CT acquisition to repeat!!!");
  auto t0 = std::chrono::steady_clock::now();
  engine_check(ggh_replay_acquisition(ctx), "acquisition");
  auto t1 = std::chrono::steady_clock::now();
  last_run_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  save_acquisition();
  log(0, "GGEMS", "GGEMS acquisition succeeded");
}

void GGEMS::save_acquisition() {
  for (GGEMSNavigator* nav : GGEMSNavigatorManager::get().navigators) {
    if (auto* ct = dynamic_cast<GGEMSCTSystem*>(nav)) ct->save_acquisition(ctx, acquisition_projections);
    else if (nav->dosimetry) nav->dosimetry->save(ctx);
  }
}

void GGEMS::autotune(uint64_t n_trial) {
  if (!initialized) raise("GGEMS", "GGEMS is not initialized!!!");
  auto& sources = GGEMSSourceManager::get().sources;
  if (sources.empty()) raise("GGEMS", "No source registered!!!");
  upload_source(*sources.front());
  ggh_launch_desc best;
  engine_check(ggh_autotune(ctx, n_trial, &best), "autotune");
  log(1, "GGEMS", "autotuned launch: round steps " + std::to_string(best.round_steps) + ", interaction batch " +
                      std::to_string(best.interact_batch) + ", navigation passes " +
                      std::to_string(best.cold_passes) + ", blocks per CU " + std::to_string(best.blocks_per_cu));
}
