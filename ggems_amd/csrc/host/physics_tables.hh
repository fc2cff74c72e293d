// physics_tables.hh -- host-side physics tables of one navigator: material
// tables, photon range cuts, cross-section tables, attenuation tables.
#pragma once

#include <string>
#include <vector>

#include "host_common.hh"

namespace gghost {

constexpr int kCompton = 0, kPhotoelectric = 1, kRayleigh = 2;

// ---- materials database (src/materials/GGEMSMaterialsDatabaseManager.cc:97-141) ----
struct DatabaseMaterial {
  float density = 0.0f;  // internal units
  std::vector<std::string> elements;
  std::vector<float> fractions;
};

class MaterialDatabase {
 public:
  static MaterialDatabase& get();
  void load(const std::string& filename);
  bool ready() const { return !materials_.empty(); }
  const DatabaseMaterial& material(const std::string& name) const;
  bool has(const std::string& name) const { return materials_.count(name) != 0; }
  std::vector<std::string> names() const;
  void clear() { materials_.clear(); }

 private:
  std::map<std::string, DatabaseMaterial> materials_;
};

struct Element {
  int z = 0;
  float molar_mass = 0.0f;  // internal units (M * g/mol)
  float mean_excitation = 0.0f;
};
const Element& element(const std::string& name);
std::vector<std::string> element_names();

// ---- material tables (src/materials/GGEMSMaterials.cc:209-300) ------------------------
struct MaterialTables {
  std::vector<std::string> names;
  std::vector<int> n_elements, element_index;
  std::vector<float> density, photon_cut;
  std::vector<uint8_t> z;
  std::vector<float> atom_density, mass_fraction;
  int total_elements() const { return (int)z.size(); }
};

MaterialTables build_material_tables(const std::vector<std::string>& names);
// GGEMSRangeCuts::ConvertToEnergy, gamma branch (src/physics/GGEMSRangeCuts.cc:129-238)
float photon_energy_cut(const MaterialTables& mt, int mat, float distance_cut, float table_min_energy);

// ---- cross sections (src/physics/GGEMSCrossSections.cc, GGEMSEMProcess.cc) -------------
std::vector<float> energy_grid(int nbins, float emin, float emax);
float cs_compton_atom(float energy, int z);
float cs_photoelectric_atom(float energy, int z);
float cs_rayleigh_atom(float energy, int z);
float cs_atom(int process, float energy, int z);

struct CrossSectionTables {
  int nbins = 0;
  float emin = 0, emax = 0;
  std::vector<int> processes;                 // activation order
  std::vector<float> energy_bins;             // [nbins]
  std::vector<float> cs;                      // [3][n_mat][nbins]
  std::vector<float> cs_atom;                 // [3][101][nbins]
};

CrossSectionTables build_cross_sections(const MaterialTables& mt, const std::vector<int>& processes, int nbins,
                                        float emin, float emax);
// GGEMSCrossSections::GetPhotonCrossSection: interpolated sigma/rho in cm2/g
float photon_cross_section(const CrossSectionTables& t, const MaterialTables& mt, int process, int mat, float energy);

// ---- attenuation (src/physics/GGEMSAttenuations.cc:142-231) -----------------------------
struct AttenuationTables {
  std::vector<float> energy_bins;  // [220]
  std::vector<float> mu, mu_en;    // [n_mat][220] cm^-1
};
AttenuationTables build_attenuations(const MaterialTables& mt);
float attenuation(const AttenuationTables& t, int mat, float energy, bool energy_absorption);

int process_id(const std::string& name);
const char* process_name(int id);

}  // namespace gghost
