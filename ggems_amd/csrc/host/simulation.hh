// simulation.hh -- GGEMS-compatible host objects behind the extern "C" API
// of include/ggems.h.  The struct names are the opaque handle types of that
// API; the behaviour (registration, units, checks, outputs) follows GGEMS,
// while the transport itself runs in the HIP engine (include/ggems_engine.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/ggems_engine.h"
#include "host_common.hh"
#include "physics_tables.hh"

struct GGEMSDosimetryCalculator;

// GGEMSMaterials: one navigator's material list (src/materials/GGEMSMaterials.cc)
struct GGEMSMaterials {
  std::vector<std::string> names;
  float photon_distance_cut = 1.e-3f;  // PHOTON_DISTANCE_CUT, 1 um (GGEMSProcessConstants.hh)
  gghost::MaterialTables tables;
  bool built = false;
  void add(const std::string& name);
  void set_distance_cut(const std::string& particle, float value, const std::string& unit);
  void initialize();  // material tables + photon cuts with the process-manager table minimum
  int index(const std::string& name) const;
};

// GGEMSCrossSections (src/physics/GGEMSCrossSections.cc)
struct GGEMSCrossSections {
  explicit GGEMSCrossSections(GGEMSMaterials* m) : materials(m) {}
  GGEMSMaterials* materials;
  std::vector<int> processes;  // activation order
  gghost::CrossSectionTables tables;
  bool built = false;
  void add_process(const std::string& process, const std::string& particle);
  void initialize();
};

// GGEMSAttenuations (src/physics/GGEMSAttenuations.cc)
struct GGEMSAttenuations {
  GGEMSAttenuations(GGEMSMaterials* m, GGEMSCrossSections* c) : materials(m), cross_sections(c) {}
  GGEMSMaterials* materials;
  GGEMSCrossSections* cross_sections;
  gghost::AttenuationTables tables;
  bool built = false;
  void initialize();
};

// GGEMSNavigator base (src/navigators/GGEMSNavigator.cc)
struct GGEMSNavigator {
  explicit GGEMSNavigator(const std::string& name);
  virtual ~GGEMSNavigator();
  std::string name;
  GGEMSMaterials materials;
  GGEMSCrossSections cross_sections{&materials};
  float position[3] = {0, 0, 0}, rotation[3] = {0, 0, 0};
  bool update_pos = false, update_rot = false;
  float threshold = 0.0f;
  std::string output_basename;
  GGEMSDosimetryCalculator* dosimetry = nullptr;
  void set_position(float x, float y, float z, const std::string& unit);
  void set_rotation(float x, float y, float z, const std::string& unit);
  virtual void initialize() = 0;                 // geometry + tables
  virtual int register_solids(ggh_context* ctx, int material_offset) = 0;  // returns solids added
  virtual void save_results(ggh_context* ctx) = 0;
  // the solid a dose map covers: its engine index, voxel size (zero for a
  // mesh) and OBB borders (GGEMSSolid::GetVoxelSizes / GetOBBGeometry);
  // false for navigators that take no dose map
  virtual bool dose_geometry(int&, float[3], float[3], float[3]) const { return false; }
  // the engine solid of a phantom navigator (voxelized or meshed), else -1
  virtual int phantom_solid() const { return -1; }
};

// GGEMSVoxelizedPhantom + GGEMSVoxelizedSolid (src/navigators/GGEMSVoxelizedPhantom.cc,
// src/geometries/GGEMSVoxelizedSolid.cc)
struct GGEMSVoxelizedPhantom : GGEMSNavigator {
  explicit GGEMSVoxelizedPhantom(const std::string& n) : GGEMSNavigator(n) {}
  std::string phantom_file, range_file;
  int dims[3] = {0, 0, 0};
  float voxel_size[3] = {0, 0, 0};
  float border_min[3] = {0, 0, 0}, border_max[3] = {0, 0, 0};
  std::vector<uint8_t> labels;
  gghost::Transform transform;
  int solid_index = -1;
  void initialize() override;
  int register_solids(ggh_context* ctx, int material_offset) override;
  void save_results(ggh_context* ctx) override;
  bool dose_geometry(int& solid, float vs[3], float lo[3], float hi[3]) const override;
  int phantom_solid() const override { return solid_index; }
};

// GGEMSMeshedPhantom + GGEMSMeshedSolid (src/navigators/GGEMSMeshedPhantom.cc,
// src/geometries/GGEMSMeshedSolid.cc): one material, triangles moved to
// global coordinates, the octree box as OBB
struct GGEMSMeshedPhantom : GGEMSNavigator {
  explicit GGEMSMeshedPhantom(const std::string& n) : GGEMSNavigator(n) {}
  std::string phantom_file;
  int octree_depth = 1;
  std::vector<float> triangles;  // [n][9], global
  float obb_min[3] = {0, 0, 0}, obb_max[3] = {0, 0, 0};
  int solid_index = -1;
  void initialize() override;
  int register_solids(ggh_context* ctx, int material_offset) override;
  void save_results(ggh_context* ctx) override;
  bool dose_geometry(int& solid, float vs[3], float lo[3], float hi[3]) const override;
  int phantom_solid() const override { return solid_index; }
};

// GGEMSCTSystem + GGEMSSystem + GGEMSSolidBox (src/navigators/GGEMSCTSystem.cc,
// GGEMSSystem.cc, src/geometries/GGEMSSolidBox.cc)
struct GGEMSCTSystem : GGEMSNavigator {
  explicit GGEMSCTSystem(const std::string& n) : GGEMSNavigator(n) {}
  std::string type;
  size_t modules[2] = {0, 0};
  size_t elements[3] = {0, 0, 0};
  float element_size[3] = {0, 0, 0};
  float sid = 0.0f, sdd = 0.0f;
  float global_position[3] = {0, 0, 0};
  bool store_scatter = false;
  std::vector<gghost::Transform> module_transforms;
  std::vector<int> solid_indices;
  void initialize() override;
  int register_solids(ggh_context* ctx, int material_offset) override;
  void save_results(ggh_context* ctx) override;
  void save_acquisition(ggh_context* ctx, int n_projections);  // image stacks [n_projections][rows][cols]
};

// GGEMSDosimetryCalculator (src/navigators/GGEMSDosimetryCalculator.cc)
struct GGEMSDosimetryCalculator {
  std::string output = "dosi";
  GGEMSNavigator* navigator = nullptr;
  float dosel_size[3] = {-1.0f, -1.0f, -1.0f};
  bool photon_tracking = false, edep = false, hit = false, edep_squared = false, uncertainty = false;
  float scale = 1.0f;
  bool water_reference = false;
  float minimum_density = 0.0f;
  int n_dosels[3] = {0, 0, 0};
  float used_size[3] = {0, 0, 0};
  void initialize(ggh_context* ctx);
  void save(ggh_context* ctx);
};

// GGEMSXRaySource + GGEMSSource (src/sources/GGEMSXRaySource.cc, GGEMSSource.cc)
struct GGEMSXRaySource {
  explicit GGEMSXRaySource(const std::string& name);
  ~GGEMSXRaySource();
  std::string name;
  uint64_t n_particles = 0;
  int particle_type = 99;
  gghost::Transform transform;
  bool position_set = false, rotation_set = false, aperture_set = false, focal_set = false;
  float aperture = 0.0f;
  float focal[3] = {0, 0, 0};
  std::vector<std::pair<double, double>> spectrum;  // (energy, intensity): GGEMSEnergyMapping
  bool is_interp = true;
  void set_position(float x, float y, float z, const std::string& unit);
  void set_rotation(float x, float y, float z, const std::string& unit);
  void set_monoenergy(float e, const std::string& unit);
  void set_polyenergy(const std::string& file);
  void set_energy_peak(float e, const std::string& unit, float intensity);
  void check() const;
  // GGEMSSource::FillEnergy (src/sources/GGEMSSource.cc:265-316)
  void fill_energy(std::vector<float>& energy, std::vector<float>& cdf) const;
};

struct GGEMSProcessesManager {
  int nbins = 220;                       // CROSS_SECTION_TABLE_NUMBER_BINS
  float emin = 990.0f * 1.e-6f;          // CROSS_SECTION_TABLE_ENERGY_MIN
  float emax = 250.0f * 1.0f;            // CROSS_SECTION_TABLE_ENERGY_MAX
  bool print_tables = false;
  static GGEMSProcessesManager& get();
};

struct GGEMSRangeCutsManager {
  static GGEMSRangeCutsManager& get();
};
struct GGEMSMaterialsDatabaseManager {
  static GGEMSMaterialsDatabaseManager& get();
};
struct GGEMSSourceManager {
  std::vector<GGEMSXRaySource*> sources;
  static GGEMSSourceManager& get();
};
struct GGEMSNavigatorManager {
  std::vector<GGEMSNavigator*> navigators;
  static GGEMSNavigatorManager& get();
  GGEMSNavigator* find(const std::string& name) const;
};

// HIP device selection shim of GGEMSOpenCLManager (src/global/GGEMSOpenCLManager.cc)
struct GGEMSOpenCLManager {
  int device = 0;
  std::vector<float> balancing;
  static GGEMSOpenCLManager& get();
  void activate(const std::string& type, const std::string& vendor);
};

struct GGEMSProfilerManager {
  static GGEMSProfilerManager& get();
};
struct GGEMSRAMManager {
  static GGEMSRAMManager& get();
};

// GGEMS (src/global/GGEMS.cc)
struct GGEMS {
  ~GGEMS();
  bool verbose_opencl = false, verbose_material = false, verbose_source = false, verbose_navigator = false,
       verbose_memory = false, verbose_process = false, verbose_range_cuts = false, verbose_random = false,
       verbose_profiling = false, verbose_tracking = false;
  int tracking_id = 0;
  uint32_t seed = 0;
  ggh_context* ctx = nullptr;
  bool initialized = false;
  double last_run_ms = 0.0;
  void initialize(uint32_t seed);
  void run();
  void upload_source(const GGEMSXRaySource& src);
  // CT acquisition (one launch, captured as a hipGraph): projection p turns
  // the source and every CT system by rotation[p] (radians) on top of their
  // placement and transports n_particles[p] primaries
  // phantom (optional): projection p sees only phantom navigator phantom[p]
  // (rank among the phantom navigators in registration order; -1: all)
  void run_acquisition(size_t n, const float* rx, const float* ry, const float* rz, const uint64_t* n_particles,
                       const int32_t* phantom = nullptr);
  void repeat_acquisition();  // replays the captured acquisition (more primaries, same placements)
  void autotune(uint64_t n_trial);
  void save_acquisition();
  int acquisition_projections = 0;
};

// the engine context of the most recently initialised GGEMS object
ggh_context* current_engine_context();
