// api.cc -- extern "C" functions of libggems.so (include/ggems.h).
//
// Each function is the one-line forwarder GGEMS exports for the same name
// (e.g. src/navigators/GGEMSCTSystem.cc:311-449, src/sources/GGEMSXRaySource.cc:282-366),
// wrapped so that an error is reported and recorded instead of unwinding
// through the C boundary.
#include <cstring>
#include <iostream>

#include "simulation.hh"
#include "volume_creator.hh"
// checks every definition below against the published declarations
#include "../../../include/ggems.h"

using namespace gghost;

#define API_TRY \
  try {         \
    clear_error();
#define API_CATCH(ret)                   \
  }                                      \
  catch (const std::exception& e) {      \
    record_error(e.what());              \
    return ret;                          \
  }

namespace {
std::string str(const char* s) { return s ? std::string(s) : std::string(); }
}  // namespace

extern "C" {

// ---- GGEMS --------------------------------------------------------------------
GGEMS* create_ggems(void) { return new (std::nothrow) GGEMS(); }
void delete_ggems(GGEMS* g) { delete g; }
void initialize_ggems(GGEMS* g, uint32_t seed) {
  API_TRY g->initialize(seed);
  API_CATCH()
}
void set_opencl_verbose_ggems(GGEMS* g, bool f) { g->verbose_opencl = f; }
void set_material_database_verbose_ggems(GGEMS* g, bool f) { g->verbose_material = f; }
void set_source_ggems(GGEMS* g, bool f) { g->verbose_source = f; }
void set_navigator_ggems(GGEMS* g, bool f) { g->verbose_navigator = f; }
void set_memory_ram_ggems(GGEMS* g, bool f) { g->verbose_memory = f; }
void set_process_ggems(GGEMS* g, bool f) { g->verbose_process = f; }
void set_range_cuts_ggems(GGEMS* g, bool f) { g->verbose_range_cuts = f; }
void set_random_ggems(GGEMS* g, bool f) { g->verbose_random = f; }
void set_profiling_ggems(GGEMS* g, bool f) { g->verbose_profiling = f; }
void set_tracking_ggems(GGEMS* g, bool f, int id) {
  g->verbose_tracking = f;
  g->tracking_id = id;
}
void run_ggems(GGEMS* g) {
  API_TRY g->run();
  API_CATCH()
}
void run_acquisition_phantoms_ggems(GGEMS* g, size_t n, const float* rx, const float* ry, const float* rz,
                                    char const* unit, const uint64_t* n_particles, const int32_t* phantoms) {
  API_TRY if (!rx || !ry || !rz || !n_particles) raise("GGEMS", "null acquisition arguments");
  std::vector<float> a(3 * n);
  for (size_t p = 0; p < n; ++p) {
    a[p] = angle_unit(rx[p], str(unit));
    a[n + p] = angle_unit(ry[p], str(unit));
    a[2 * n + p] = angle_unit(rz[p], str(unit));
  }
  g->run_acquisition(n, a.data(), a.data() + n, a.data() + 2 * n, n_particles, phantoms);
  API_CATCH()
}
void run_acquisition_ggems(GGEMS* g, size_t n, const float* rx, const float* ry, const float* rz, char const* unit,
                           const uint64_t* n_particles) {
  run_acquisition_phantoms_ggems(g, n, rx, ry, rz, unit, n_particles, nullptr);
}
void repeat_acquisition_ggems(GGEMS* g) {
  API_TRY g->repeat_acquisition();
  API_CATCH()
}
void autotune_ggems(GGEMS* g, uint64_t n_trial) {
  API_TRY g->autotune(n_trial);
  API_CATCH()
}

// ---- OpenCL manager (HIP device selection) --------------------------------------------
GGEMSOpenCLManager* get_instance_ggems_opencl_manager(void) { return &GGEMSOpenCLManager::get(); }
void print_infos_opencl_manager(GGEMSOpenCLManager* m) {
  API_TRY int n = 0;
  ggh_device_count(&n);
  std::cout << "[GGEMS HIP] devices: " << n << ", activated: " << m->device << std::endl;
  API_CATCH()
}
void set_device_index_ggems_opencl_manager(GGEMSOpenCLManager* m, size_t id) {
  API_TRY m->activate(std::to_string(id), "");
  API_CATCH()
}
void set_device_to_activate_opencl_manager(GGEMSOpenCLManager* m, char const* type, char const* vendor) {
  API_TRY m->activate(str(type), str(vendor));
  API_CATCH()
}
void clean_opencl_manager(GGEMSOpenCLManager*) {}
void set_device_balancing_opencl_manager(GGEMSOpenCLManager* m, char const* b) {
  API_TRY m->balancing.clear();
  std::string s = str(b), tok;
  size_t p = 0;
  while (p <= s.size()) {
    size_t q = s.find(';', p);
    tok = s.substr(p, q == std::string::npos ? std::string::npos : q - p);
    if (!tok.empty()) m->balancing.push_back(std::stof(tok));
    if (q == std::string::npos) break;
    p = q + 1;
  }
  float sum = 0.0f;
  for (float v : m->balancing) sum += v;
  if (!m->balancing.empty() && (sum < 0.999f || sum > 1.001f))
    raise("GGEMSOpenCLManager", "Device balancing has to be 1 !!!");
  API_CATCH()
}

// ---- materials -------------------------------------------------------------------------
GGEMSMaterials* create_ggems_materials(void) { return new (std::nothrow) GGEMSMaterials(); }
void add_material_ggems_materials(GGEMSMaterials* m, char const* name) {
  API_TRY m->add(str(name));
  API_CATCH()
}
void initialize_ggems_materials(GGEMSMaterials* m) {
  API_TRY m->initialize();
  API_CATCH()
}
void print_material_properties_ggems_materials(GGEMSMaterials* m) {
  API_TRY if (!m->built) m->initialize();
  for (size_t i = 0; i < m->names.size(); ++i)
    std::cout << "[GGEMS Materials] " << m->names[i] << ": density " << m->tables.density[i] / (unit::g / unit::cm3)
              << " g/cm3, elements " << m->tables.n_elements[i] << ", photon cut "
              << m->tables.photon_cut[i] / unit::keV << " keV" << std::endl;
  API_CATCH()
}
float get_density_ggems_materials(GGEMSMaterials* m, char const* name) {
  API_TRY if (!m->built) m->initialize();
  return m->tables.density[m->index(str(name))] / unit::g * unit::cm3;
  API_CATCH(0.0f)
}
// GGEMSMaterials::GetEnergyCut: the cut for a given distance (src/materials/GGEMSMaterials.cc:345-370)
float get_energy_cut_ggems_materials(GGEMSMaterials* m, char const* name, char const* particle, float distance,
                                     char const* unit_) {
  API_TRY if (!m->built) m->initialize();
  if (str(particle) != "gamma") raise("GGEMSMaterials", "only the photon ('gamma') cut is computed");
  float d = distance_unit(distance, str(unit_));
  return photon_energy_cut(m->tables, m->index(str(name)), d, GGEMSProcessesManager::get().emin) / unit::keV;
  API_CATCH(0.0f)
}
float get_atomic_number_density_ggems_materials(GGEMSMaterials* m, char const* name) {
  API_TRY if (!m->built) m->initialize();
  // GGEMSMaterials::GetAtomicNumberDensity (src/materials/GGEMSMaterials.cc:334-350) reads the
  // per-element array at the material's index; kept as is
  int i = m->index(str(name));
  return m->tables.atom_density[(size_t)i] / (unit::mol / unit::cm3);
  API_CATCH(0.0f)
}
float get_photon_cut_ggems_materials(GGEMSMaterials* m, char const* name) {
  API_TRY if (!m->built) m->initialize();
  return m->tables.photon_cut[m->index(str(name))];
  API_CATCH(0.0f)
}
void clean_ggems_materials(GGEMSMaterials* m) { delete m; }

// ---- materials database ---------------------------------------------------------------
GGEMSMaterialsDatabaseManager* get_instance_materials_manager(void) { return &GGEMSMaterialsDatabaseManager::get(); }
void set_materials_ggems_materials_manager(GGEMSMaterialsDatabaseManager*, char const* filename) {
  API_TRY MaterialDatabase::get().load(str(filename));
  API_CATCH()
}
void print_available_chemical_elements_ggems_materials_manager(GGEMSMaterialsDatabaseManager*) {
  API_TRY for (const std::string& n : element_names()) std::cout << "    * " << n << std::endl;
  API_CATCH()
}
void print_available_materials_ggems_materials_manager(GGEMSMaterialsDatabaseManager*) {
  API_TRY for (const std::string& n : MaterialDatabase::get().names()) std::cout << "    * " << n << std::endl;
  API_CATCH()
}

// ---- range cuts / processes ----------------------------------------------------------------
GGEMSRangeCutsManager* get_instance_range_cuts_manager(void) { return &GGEMSRangeCutsManager::get(); }
void set_cut_range_cuts_manager(GGEMSRangeCutsManager*, char const* phantom, char const* particle, float value,
                                char const* unit_) {
  API_TRY std::string p = str(phantom);
  if (p == "all") {
    for (GGEMSNavigator* n : GGEMSNavigatorManager::get().navigators)
      n->materials.set_distance_cut(str(particle), value, str(unit_));
  } else {
    GGEMSNavigatorManager::get().find(p)->materials.set_distance_cut(str(particle), value, str(unit_));
  }
  API_CATCH()
}

GGEMSProcessesManager* get_instance_processes_manager(void) { return &GGEMSProcessesManager::get(); }
void add_process_processes_manager(GGEMSProcessesManager*, char const* process, char const* particle,
                                   char const* phantom) {
  API_TRY std::string p = str(phantom);
  if (p == "all") {
    for (GGEMSNavigator* n : GGEMSNavigatorManager::get().navigators)
      n->cross_sections.add_process(str(process), str(particle));
  } else {
    GGEMSNavigatorManager::get().find(p)->cross_sections.add_process(str(process), str(particle));
  }
  API_CATCH()
}
void set_cross_section_table_number_of_bins_processes_manager(GGEMSProcessesManager* m, size_t n) {
  API_TRY if (n > 2048) {
    log(0, "GGEMSProcessesManager", "Number of bins in the cross section table > 2048, set to 2048.");
    n = 2048;
  }
  if (n < 2) raise("GGEMSProcessesManager", "the cross section table needs at least 2 bins");
  m->nbins = (int)n;
  API_CATCH()
}
void set_cross_section_table_minimum_energy_processes_manager(GGEMSProcessesManager* m, float e, char const* u) {
  API_TRY float v = energy_unit(e, str(u));
  if (v < 990.0f * 1.e-6f)
    raise("GGEMSProcessesManager", "The minimum of energy in the cross section table is 990 eV");
  m->emin = v;
  API_CATCH()
}
void set_cross_section_table_maximum_energy_processes_manager(GGEMSProcessesManager* m, float e, char const* u) {
  API_TRY float v = energy_unit(e, str(u));
  if (v > 250.0f) raise("GGEMSProcessesManager", "The maximum of energy in the cross section table is 250 MeV");
  m->emax = v;
  API_CATCH()
}
void print_infos_processes_manager(GGEMSProcessesManager* m) {
  std::cout << "[GGEMS Processes] table: " << m->nbins << " bins, " << m->emin << " - " << m->emax << " MeV"
            << std::endl;
  for (GGEMSNavigator* n : GGEMSNavigatorManager::get().navigators) {
    std::cout << "    " << n->name << ":";
    for (int p : n->cross_sections.processes) std::cout << " " << process_name(p);
    std::cout << std::endl;
  }
}
void print_available_processes_manager(GGEMSProcessesManager*) {
  std::cout << "Available processes: 'Compton', 'Photoelectric', 'Rayleigh' (gamma)" << std::endl;
}
void print_tables_processes_manager(GGEMSProcessesManager* m, bool f) { m->print_tables = f; }

// ---- cross sections / attenuations ------------------------------------------------------
GGEMSCrossSections* create_ggems_cross_sections(GGEMSMaterials* m) { return new (std::nothrow) GGEMSCrossSections(m); }
void add_process_ggems_cross_sections(GGEMSCrossSections* c, char const* process, char const* particle, bool) {
  API_TRY c->add_process(str(process), str(particle));
  API_CATCH()
}
void initialize_ggems_cross_sections(GGEMSCrossSections* c) {
  API_TRY c->initialize();
  API_CATCH()
}
float get_cs_ggems_cross_sections(GGEMSCrossSections* c, char const* process, char const* material, float e,
                                  char const* u) {
  API_TRY if (!c->built) c->initialize();
  int p = process_id(str(process));
  return photon_cross_section(c->tables, c->materials->tables, p, c->materials->index(str(material)),
                              energy_unit(e, str(u)));
  API_CATCH(0.0f)
}
void get_cross_section_table_ggems_cross_sections(GGEMSCrossSections* c, char const* process, char const* material,
                                                  float* out, size_t nbins) {
  API_TRY if (!c->built) c->initialize();
  if ((int)nbins != c->tables.nbins) raise("GGEMSCrossSections", "nbins does not match the table");
  int p = process_id(str(process));
  int m = c->materials->index(str(material));
  const size_t nm = c->materials->names.size();
  std::memcpy(out, &c->tables.cs[((size_t)p * nm + m) * nbins], nbins * sizeof(float));
  API_CATCH()
}
void get_energy_bins_ggems_cross_sections(GGEMSCrossSections* c, float* out, size_t nbins) {
  API_TRY if (!c->built) c->initialize();
  if ((int)nbins != c->tables.nbins) raise("GGEMSCrossSections", "nbins does not match the table");
  std::memcpy(out, c->tables.energy_bins.data(), nbins * sizeof(float));
  API_CATCH()
}
void clean_ggems_cross_sections(GGEMSCrossSections* c) { delete c; }

GGEMSAttenuations* create_ggems_attenuations(GGEMSMaterials* m, GGEMSCrossSections* c) {
  return new (std::nothrow) GGEMSAttenuations(m, c);
}
void initialize_ggems_attenuations(GGEMSAttenuations* a) {
  API_TRY a->initialize();
  API_CATCH()
}
float get_mu_ggems_attenuations(GGEMSAttenuations* a, char const* material, float e, char const* u) {
  API_TRY if (!a->built) a->initialize();
  return attenuation(a->tables, a->materials->index(str(material)), energy_unit(e, str(u)), false);
  API_CATCH(0.0f)
}
float get_mu_en_ggems_attenuations(GGEMSAttenuations* a, char const* material, float e, char const* u) {
  API_TRY if (!a->built) a->initialize();
  return attenuation(a->tables, a->materials->index(str(material)), energy_unit(e, str(u)), true);
  API_CATCH(0.0f)
}
void clean_ggems_attenuations(GGEMSAttenuations* a) { delete a; }

// ---- sources ---------------------------------------------------------------------------------
GGEMSSourceManager* get_instance_ggems_source_manager(void) { return &GGEMSSourceManager::get(); }
void initialize_source_manager(GGEMSSourceManager* m, uint32_t) {
  API_TRY for (GGEMSXRaySource* s : m->sources) s->check();
  API_CATCH()
}
void print_infos_source_manager(GGEMSSourceManager* m) {
  for (GGEMSXRaySource* s : m->sources)
    std::cout << "[GGEMS Source] " << s->name << ": " << s->n_particles << " particles, " << s->spectrum.size()
              << " energy lines" << std::endl;
}
void clean_source_manager(GGEMSSourceManager*) {}

GGEMSXRaySource* create_ggems_xray_source(char const* name) { return new (std::nothrow) GGEMSXRaySource(str(name)); }
void set_position_ggems_xray_source(GGEMSXRaySource* s, float x, float y, float z, char const* u) {
  API_TRY s->set_position(x, y, z, str(u));
  API_CATCH()
}
void set_number_of_particles_xray_source(GGEMSXRaySource* s, size_t n) { s->n_particles = n; }
void set_source_particle_type_ggems_xray_source(GGEMSXRaySource* s, char const* p) {
  API_TRY std::string t = str(p);
  if (t == "gamma") s->particle_type = 0;
  else if (t == "e-") s->particle_type = 1;
  else if (t == "e+") s->particle_type = 2;
  else raise("GGEMSSourceManager", "Unknown particle!!!");
  API_CATCH()
}
void set_beam_aperture_ggems_xray_source(GGEMSXRaySource* s, float a, char const* u) {
  API_TRY s->aperture = angle_unit(a, str(u));
  s->aperture_set = true;
  API_CATCH()
}
void set_focal_spot_size_ggems_xray_source(GGEMSXRaySource* s, float w, float h, float d, char const* u) {
  API_TRY s->focal[0] = distance_unit(w, str(u));
  s->focal[1] = distance_unit(h, str(u));
  s->focal[2] = distance_unit(d, str(u));
  s->focal_set = true;
  API_CATCH()
}
void set_rotation_ggems_xray_source(GGEMSXRaySource* s, float x, float y, float z, char const* u) {
  API_TRY s->set_rotation(x, y, z, str(u));
  API_CATCH()
}
void set_monoenergy_ggems_xray_source(GGEMSXRaySource* s, float e, char const* u) {
  API_TRY s->set_monoenergy(e, str(u));
  API_CATCH()
}
void set_polyenergy_ggems_xray_source(GGEMSXRaySource* s, char const* f) {
  API_TRY s->set_polyenergy(str(f));
  API_CATCH()
}
void set_energy_peak_ggems_xray_source(GGEMSXRaySource* s, float e, char const* u, float intensity) {
  API_TRY s->set_energy_peak(e, str(u), intensity);
  API_CATCH()
}

// ---- verbosity, profiler, RAM ----------------------------------------------------------------
void set_ggems_verbose(int v) { set_verbosity(v); }
GGEMSProfilerManager* get_instance_profiler_manager(void) { return &GGEMSProfilerManager::get(); }
void print_summary_profiler_manager(GGEMSProfilerManager*) {
  float ms = 0.0f;
  if (ggh_context* c = current_engine_context()) ggh_last_kernel_ms(c, &ms);
  std::cout << "[GGEMS Profiler] transport kernel: " << ms << " ms" << std::endl;
}
GGEMSRAMManager* get_instance_ggems_ram_manager(void) { return &GGEMSRAMManager::get(); }
void print_infos_ram_manager(GGEMSRAMManager*) {}

// ---- voxelized phantom -------------------------------------------------------------------------
GGEMSVoxelizedPhantom* create_ggems_voxelized_phantom(char const* name) {
  return new (std::nothrow) GGEMSVoxelizedPhantom(str(name));
}
void set_phantom_file_ggems_voxelized_phantom(GGEMSVoxelizedPhantom* p, char const* f, char const* r) {
  p->phantom_file = str(f);
  p->range_file = str(r);
}
void set_position_ggems_voxelized_phantom(GGEMSVoxelizedPhantom* p, float x, float y, float z, char const* u) {
  API_TRY p->set_position(x, y, z, str(u));
  API_CATCH()
}
void set_rotation_ggems_voxelized_phantom(GGEMSVoxelizedPhantom* p, float x, float y, float z, char const* u) {
  API_TRY p->set_rotation(x, y, z, str(u));
  API_CATCH()
}
// OpenGL display settings: accepted; this library has no display
void set_visible_ggems_voxelized_phantom(GGEMSVoxelizedPhantom*, bool) {}
void set_material_visible_ggems_voxelized_phantom(GGEMSVoxelizedPhantom*, char const*, bool) {}
void set_material_color_ggems_voxelized_phantom(GGEMSVoxelizedPhantom*, char const*, unsigned char, unsigned char,
                                                unsigned char) {}
void set_material_color_name_ggems_voxelized_phantom(GGEMSVoxelizedPhantom*, char const*, char const*) {}

// ---- meshed phantom ---------------------------------------------------------------------------
GGEMSMeshedPhantom* create_ggems_meshed_phantom(char const* name) {
  return new (std::nothrow) GGEMSMeshedPhantom(str(name));
}
void set_phantom_file_ggems_meshed_phantom(GGEMSMeshedPhantom* p, char const* f) { p->phantom_file = str(f); }
void set_position_ggems_meshed_phantom(GGEMSMeshedPhantom* p, float x, float y, float z, char const* u) {
  API_TRY p->set_position(x, y, z, str(u));
  API_CATCH()
}
void set_rotation_ggems_meshed_phantom(GGEMSMeshedPhantom* p, float x, float y, float z, char const* u) {
  API_TRY p->set_rotation(x, y, z, str(u));
  API_CATCH()
}
void set_material_name_ggems_meshed_phantom(GGEMSMeshedPhantom* p, char const* material) {
  // GGEMSNavigator::SetMaterialName (src/navigators/GGEMSNavigator.cc:258-261)
  API_TRY p->materials.add(str(material));
  API_CATCH()
}
void set_mesh_octree_depth_ggems_meshed_phantom(GGEMSMeshedPhantom* p, int depth) { p->octree_depth = depth; }
// OpenGL display settings: accepted; this library has no display
void set_visible_ggems_meshed_phantom(GGEMSMeshedPhantom*, bool) {}
void set_material_color_ggems_meshed_phantom(GGEMSMeshedPhantom*, char const*, unsigned char, unsigned char,
                                             unsigned char) {}
void set_material_color_name_ggems_meshed_phantom(GGEMSMeshedPhantom*, char const*, char const*) {}

// ---- CT system -----------------------------------------------------------------------------------
GGEMSCTSystem* create_ggems_ct_system(char const* name) { return new (std::nothrow) GGEMSCTSystem(str(name)); }
void set_number_of_modules_ggems_ct_system(GGEMSCTSystem* c, size_t x, size_t y) {
  c->modules[0] = x;
  c->modules[1] = y;
}
void set_ct_system_type_ggems_ct_system(GGEMSCTSystem* c, char const* t) {
  API_TRY std::string s = str(t);
  for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
  if (s != "curved" && s != "flat") raise("GGEMSCTSystem", "Available CT system types: 'curved' or 'flat'");
  c->type = s;
  API_CATCH()
}
void set_number_of_detection_elements_ggems_ct_system(GGEMSCTSystem* c, size_t x, size_t y, size_t z) {
  c->elements[0] = x;
  c->elements[1] = y;
  c->elements[2] = z;
}
void set_size_of_detection_elements_ggems_ct_system(GGEMSCTSystem* c, float x, float y, float z, char const* u) {
  API_TRY c->element_size[0] = distance_unit(x, str(u));
  c->element_size[1] = distance_unit(y, str(u));
  c->element_size[2] = distance_unit(z, str(u));
  API_CATCH()
}
void set_material_name_ggems_ct_system(GGEMSCTSystem* c, char const* m) {
  API_TRY c->materials.add(str(m));
  API_CATCH()
}
void set_source_isocenter_distance_ggems_ct_system(GGEMSCTSystem* c, float d, char const* u) {
  API_TRY c->sid = distance_unit(d, str(u));
  API_CATCH()
}
void set_source_detector_distance_ggems_ct_system(GGEMSCTSystem* c, float d, char const* u) {
  API_TRY c->sdd = distance_unit(d, str(u));
  API_CATCH()
}
void set_rotation_ggems_ct_system(GGEMSCTSystem* c, float x, float y, float z, char const* u) {
  API_TRY c->set_rotation(x, y, z, str(u));
  API_CATCH()
}
void set_threshold_ggems_ct_system(GGEMSCTSystem* c, float t, char const* u) {
  API_TRY c->threshold = energy_unit(t, str(u));
  API_CATCH()
}
void set_save_ggems_ct_system(GGEMSCTSystem* c, char const* b) { c->output_basename = str(b); }
void store_scatter_ggems_ct_system(GGEMSCTSystem* c, bool s) { c->store_scatter = s; }
void set_visible_ggems_ct_system(GGEMSCTSystem*, bool) {}
void set_material_color_ggems_ct_system(GGEMSCTSystem*, char const*, unsigned char, unsigned char, unsigned char) {}
void set_material_color_name_ggems_ct_system(GGEMSCTSystem*, char const*, char const*) {}
void set_global_system_position_ggems_ct_system(GGEMSCTSystem* c, float x, float y, float z, char const* u) {
  API_TRY c->global_position[0] = distance_unit(x, str(u));
  c->global_position[1] = distance_unit(y, str(u));
  c->global_position[2] = distance_unit(z, str(u));
  API_CATCH()
}

// ---- dosimetry ---------------------------------------------------------------------------------------
GGEMSDosimetryCalculator* create_ggems_dosimetry_calculator(void) {
  return new (std::nothrow) GGEMSDosimetryCalculator();
}
void delete_dosimetry_calculator(GGEMSDosimetryCalculator* d) {
  for (GGEMSNavigator* n : GGEMSNavigatorManager::get().navigators)
    if (n->dosimetry == d) n->dosimetry = nullptr;
  delete d;
}
void scale_factor_dosimetry_calculator(GGEMSDosimetryCalculator* d, float s) { d->scale = s; }
void water_reference_dosimetry_calculator(GGEMSDosimetryCalculator* d, bool f) { d->water_reference = f; }
void minimum_density_dosimetry_calculator(GGEMSDosimetryCalculator* d, float v, char const* u) {
  API_TRY d->minimum_density = density_unit(v, str(u));
  API_CATCH()
}
void set_dosel_size_dosimetry_calculator(GGEMSDosimetryCalculator* d, float x, float y, float z, char const* u) {
  API_TRY d->dosel_size[0] = distance_unit(x, str(u));
  d->dosel_size[1] = distance_unit(y, str(u));
  d->dosel_size[2] = distance_unit(z, str(u));
  API_CATCH()
}
void set_dose_output_dosimetry_calculator(GGEMSDosimetryCalculator* d, char const* f) { d->output = str(f); }
void dose_photon_tracking_dosimetry_calculator(GGEMSDosimetryCalculator* d, bool f) { d->photon_tracking = f; }
void dose_edep_dosimetry_calculator(GGEMSDosimetryCalculator* d, bool f) { d->edep = f; }
void dose_hit_dosimetry_calculator(GGEMSDosimetryCalculator* d, bool f) { d->hit = f; }
void dose_edep_squared_dosimetry_calculator(GGEMSDosimetryCalculator* d, bool f) { d->edep_squared = f; }
void dose_uncertainty_dosimetry_calculator(GGEMSDosimetryCalculator* d, bool f) { d->uncertainty = f; }
void dose_tle_navigator(GGEMSDosimetryCalculator*, bool f) {
  API_TRY if (f) raise("GGEMSDosimetryCalculator", "track-length estimator (TLE) dose is not available in this library");
  API_CATCH()
}
void attach_to_navigator_dosimetry_calculator(GGEMSDosimetryCalculator* d, char const* name) {
  API_TRY GGEMSNavigator* n = GGEMSNavigatorManager::get().find(str(name));
  int solid;
  float vs[3], lo[3], hi[3];
  if (!n || !n->dose_geometry(solid, vs, lo, hi))
    raise("GGEMSDosimetryCalculator", "dosimetry attaches to a voxelized or meshed phantom");
  d->navigator = n;
  n->dosimetry = d;
  API_CATCH()
}

// ---- volume creator ------------------------------------------------------------------------------
GGEMSVolumeCreatorManager* get_instance_volume_creator_manager(void) { return &GGEMSVolumeCreatorManager::get(); }
void set_volume_dimension_volume_creator_manager(GGEMSVolumeCreatorManager* m, size_t w, size_t h, size_t d) {
  m->set_dimensions(w, h, d);
}
void set_element_sizes_volume_creator_manager(GGEMSVolumeCreatorManager* m, float x, float y, float z, char const* u) {
  API_TRY m->set_element_sizes(x, y, z, str(u));
  API_CATCH()
}
void set_output_image_filename_volume_creator_manager(GGEMSVolumeCreatorManager* m, char const* f) {
  m->output_image = str(f);
}
void set_output_range_to_material_filename_volume_creator_manager(GGEMSVolumeCreatorManager* m, char const* f) {
  m->output_range = str(f);
}
void initialize_volume_creator_manager(GGEMSVolumeCreatorManager* m) {
  API_TRY m->initialize();
  API_CATCH()
}
void write_volume_creator_manager(GGEMSVolumeCreatorManager* m) {
  API_TRY m->write();
  API_CATCH()
}
void set_material_volume_creator_manager(GGEMSVolumeCreatorManager* m, char const* material) {
  m->set_material(str(material));
}
void set_data_type_volume_creator_manager(GGEMSVolumeCreatorManager* m, char const* t) {
  API_TRY m->set_data_type(str(t));
  API_CATCH()
}
void clean_volume_creator_manager(GGEMSVolumeCreatorManager* m) { m->clean(); }
GGEMSBox* create_box(float w, float h, float d, char const* u) {
  API_TRY return new GGEMSBox(w, h, d, str(u));
  API_CATCH(nullptr)
}
GGEMSTube* create_tube(float rx, float ry, float h, char const* u) {
  API_TRY return new GGEMSTube(rx, ry, h, str(u));
  API_CATCH(nullptr)
}
GGEMSSphere* create_sphere(float r, char const* u) {
  API_TRY return new GGEMSSphere(r, str(u));
  API_CATCH(nullptr)
}
void delete_box(GGEMSBox* v) { delete v; }
void set_position_box(GGEMSBox* v, float x, float y, float z, char const* u) {
  API_TRY v->set_position(x, y, z, str(u));
  API_CATCH()
}
void set_material_box(GGEMSBox* v, char const* material) {
  API_TRY v->set_material(str(material));
  API_CATCH()
}
void set_label_value_box(GGEMSBox* v, float label) { v->set_label(label); }
void initialize_box(GGEMSBox* v) {
  API_TRY v->initialize();
  API_CATCH()
}
void draw_box(GGEMSBox* v) {
  API_TRY v->draw();
  API_CATCH()
}
void delete_tube(GGEMSTube* v) { delete v; }
void set_position_tube(GGEMSTube* v, float x, float y, float z, char const* u) {
  API_TRY v->set_position(x, y, z, str(u));
  API_CATCH()
}
void set_material_tube(GGEMSTube* v, char const* material) {
  API_TRY v->set_material(str(material));
  API_CATCH()
}
void set_label_value_tube(GGEMSTube* v, float label) { v->set_label(label); }
void initialize_tube(GGEMSTube* v) {
  API_TRY v->initialize();
  API_CATCH()
}
void draw_tube(GGEMSTube* v) {
  API_TRY v->draw();
  API_CATCH()
}
void delete_sphere(GGEMSSphere* v) { delete v; }
void set_position_sphere(GGEMSSphere* v, float x, float y, float z, char const* u) {
  API_TRY v->set_position(x, y, z, str(u));
  API_CATCH()
}
void set_material_sphere(GGEMSSphere* v, char const* material) {
  API_TRY v->set_material(str(material));
  API_CATCH()
}
void set_label_value_sphere(GGEMSSphere* v, float label) { v->set_label(label); }
void initialize_sphere(GGEMSSphere* v) {
  API_TRY v->initialize();
  API_CATCH()
}
void draw_sphere(GGEMSSphere* v) {
  API_TRY v->draw();
  API_CATCH()
}

// ---- additions ---------------------------------------------------------------------------------------
char const* get_last_error_ggems(void) {
  // read-and-clear, so that calls which cannot fail never leave a stale message behind
  thread_local std::string reported;
  reported = last_error();
  clear_error();
  return reported.c_str();
}

// GGEMS's clean_safely (python_module/__init__.py:121-124) tears the singletons down at exit;
// this also deletes every registered navigator and source so a process can set up a new scene
void clean_safely_ggems(void) {
  auto navs = GGEMSNavigatorManager::get().navigators;
  for (GGEMSNavigator* n : navs) delete n;
  auto srcs = GGEMSSourceManager::get().sources;
  for (GGEMSXRaySource* s : srcs) delete s;
  GGEMSProcessesManager& pm = GGEMSProcessesManager::get();
  pm = GGEMSProcessesManager();
  clear_error();
}
double get_run_time_ms_ggems(GGEMS* g) { return g->last_run_ms; }
void* get_engine_context_ggems(GGEMS* g) { return g->ctx; }
void get_engine_stats_ggems(GGEMS* g, uint64_t* out) {
  API_TRY if (!g->ctx) raise("GGEMS", "GGEMS is not initialized!!!");
  ggh_stats s;
  if (ggh_get_stats(g->ctx, &s) < 0) raise("GGEMSEngine", ggh_last_error());
  out[0] = s.primaries;
  out[1] = s.voxel_steps;
  out[2] = s.interactions;
  out[3] = s.detector_steps;
  out[4] = s.detector_interactions;
  out[5] = s.hops;
  out[6] = s.mesh_steps;
  out[7] = s.mesh_interactions;
  API_CATCH()
}

}  // extern "C"
