// physics_tables.cc -- host-side physics tables.
//
// The arithmetic follows the GGEMS host builders in their FP32 operation
// order (file:line per function); the library is compiled with
// -ffp-contract=off so that every table is bit-identical to a host restatement.
#include "physics_tables.hh"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

namespace gghost {

namespace {
constexpr float kAvogadro = 6.02214179e+23f;     // GGEMSConstants.hh
constexpr float kElectronMassC2 = 0.510998910f;
constexpr float kAtomicMassUnit = 0.01036426891f;

// GGEMSTextReader::RemoveSpace (src/io/GGEMSTextReader.cc:112-127)
void remove_space(std::string& s) {
  s.erase(std::remove_if(s.begin(), s.end(),
                         [](char c) { return c == '\t' || c == ' ' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }),
          s.end());
}

bool blank(const std::string& s) { return s.find_first_not_of("\t\n ") == std::string::npos; }

float parse_float(const std::string& s) { return std::strtof(s.c_str(), nullptr); }
}  // namespace

// ---------------------------------------------------------------------------
// materials database
// ---------------------------------------------------------------------------
MaterialDatabase& MaterialDatabase::get() {
  static MaterialDatabase db;
  return db;
}

void MaterialDatabase::load(const std::string& filename) {
  if (!materials_.empty()) {
    log(0, "GGEMSMaterialsDatabaseManager", "Material database if already loaded!!!");
    return;
  }
  std::ifstream in(filename);
  if (!in) raise("GGEMSMaterialsDatabaseManager", "Problem reading filename: " + filename);
  std::string line;
  while (std::getline(in, line)) {
    // SkipComment (src/io/GGEMSTextReader.cc:89-98)
    for (;;) {
      size_t p = line.find_first_not_of("\t ");
      if (p == std::string::npos || line[p] != '#') break;
      if (!std::getline(in, line)) {
        line.clear();
        break;
      }
    }
    if (blank(line)) continue;
    remove_space(line);
    DatabaseMaterial m;
    // ReadMaterialName / ReadMaterialDensity / ReadMaterialNumberOfElements (:43-89)
    std::string name = line.substr(0, line.find(':'));
    size_t d = line.find("d=");
    size_t first = line.find_first_of("0123456789", d == std::string::npos ? 0 : d);
    size_t last = line.find_first_not_of("0123456789.", first);
    float density = parse_float(line.substr(first, last == std::string::npos ? std::string::npos : last - first));
    size_t semi = line.find_first_of(';');
    std::string u = line.substr(last, semi == std::string::npos ? std::string::npos : semi - last);
    if (u == "g/cm3") density *= unit::g / unit::cm3;
    else if (u == "mg/cm3") density *= unit::mg / unit::cm3;
    else raise("GGEMSMaterialReader", "Unknown density unit in material database file!!!");
    m.density = density;
    size_t np = line.find("n=");
    size_t nf = line.find_first_of("0123456789", np == std::string::npos ? 0 : np);
    int n = std::atoi(line.c_str() + nf);
    for (int i = 0; i < n; ++i) {
      if (!std::getline(in, line)) raise("GGEMSMaterialsDatabaseManager", "truncated material " + name);
      remove_space(line);
      size_t a = line.find("name=") + 5;
      size_t b = line.find_first_of(';');
      m.elements.push_back(line.substr(a, b == std::string::npos ? std::string::npos : b - a));
      size_t f = line.find_first_of("0123456789", line.find("f="));
      m.fractions.push_back(parse_float(line.substr(f)));
    }
    materials_.insert(std::make_pair(name, m));  // the first definition of a name wins
  }
}

const DatabaseMaterial& MaterialDatabase::material(const std::string& name) const {
  auto it = materials_.find(name);
  if (it == materials_.end()) raise("GGEMSMaterialsDatabaseManager", "Material '" + name + "' not found in the database!!!");
  return it->second;
}

std::vector<std::string> MaterialDatabase::names() const {
  std::vector<std::string> out;
  for (auto& kv : materials_) out.push_back(kv.first);
  return out;
}

// chemical elements: GGEMSMaterialsDatabaseManager::AddChemicalElements
// (src/materials/GGEMSMaterialsDatabaseManager.cc:337-475), data from the blob
namespace {
std::map<std::string, Element>& element_table() {
  static std::map<std::string, Element> table;
  if (table.empty()) {
    const PhysicsData& pd = PhysicsData::get();
    size_t n = 0, nb = 0;
    const int32_t* z = pd.i32("element_z", &n);
    const float* mm = pd.f32("element_molar_mass");
    const float* ie = pd.f32("element_mean_excitation");
    const char* names = pd.bytes("element_names", &nb);
    std::istringstream ss(std::string(names, names + strnlen(names, nb)));
    std::string nm;
    for (size_t i = 0; i < n && std::getline(ss, nm); ++i) {
      Element e;
      e.z = z[i];
      e.molar_mass = mm[i] * (unit::g / unit::mol);
      e.mean_excitation = ie[i] * unit::eV;
      table[nm] = e;
    }
  }
  return table;
}
}  // namespace

const Element& element(const std::string& name) {
  auto& t = element_table();
  auto it = t.find(name);
  if (it == t.end()) raise("GGEMSMaterialsDatabaseManager", "Chemical element '" + name + "' not found!!!");
  return it->second;
}

std::vector<std::string> element_names() {
  std::vector<std::string> out;
  for (auto& kv : element_table()) out.push_back(kv.first);
  return out;
}

// ---------------------------------------------------------------------------
// material tables
// ---------------------------------------------------------------------------
MaterialTables build_material_tables(const std::vector<std::string>& names) {
  if (names.size() > 255) raise("GGEMSMaterials", "Limit of material reached. The limit is 255 materials!!!");
  const MaterialDatabase& db = MaterialDatabase::get();
  MaterialTables t;
  t.names = names;
  int idx = 0;
  for (const std::string& nm : names) {
    const DatabaseMaterial& m = db.material(nm);
    t.n_elements.push_back((int)m.elements.size());
    t.density.push_back(m.density);
    t.element_index.push_back(idx);
    t.photon_cut.push_back(0.0f);
    for (size_t j = 0; j < m.elements.size(); ++j) {
      const Element& e = element(m.elements[j]);
      t.z.push_back((uint8_t)e.z);
      t.mass_fraction.push_back(m.fractions[j]);
      // GetAtomicNumberDensity (GGEMSMaterialsDatabaseManager.hh:244-255): FP64, stored FP32
      double v = static_cast<double>(kAvogadro) / static_cast<double>(e.molar_mass) *
                 static_cast<double>(m.density) * static_cast<double>(m.fractions[j]);
      t.atom_density.push_back(static_cast<float>(v));
    }
    idx += (int)m.elements.size();
  }
  return t;
}

// ---------------------------------------------------------------------------
// photon range cut (src/physics/GGEMSRangeCuts.cc, GGEMSLogEnergyTable.cc)
// ---------------------------------------------------------------------------
namespace {
struct LogTable {
  float bin_width = 0, base_bin = 0;
  int nodes = 0;
  std::vector<float> bins, data;
  LogTable(float lo, float hi, int nbins) {
    bin_width = logf(hi / lo) / static_cast<float>(nbins);
    base_bin = logf(lo) / bin_width;
    nodes = nbins + 1;
    bins.assign(nodes, 0.0f);
    data.assign(nodes, 0.0f);
    bins[0] = lo;
    for (int i = 1; i < nodes - 1; ++i) bins[i] = expf((base_bin + static_cast<float>(i)) * bin_width);
    bins[nodes - 1] = hi;
  }
  static float interp(float xa, float ya, float xb, float yb, float x) {
    if (xa > xb) return yb;
    if (xa >= x) return ya;
    if (xb <= x) return yb;
    return ya + (x - xa) * (yb - ya) / (xb - xa);
  }
  size_t bin_location(float e) const {
    size_t bin = static_cast<size_t>(logf(e) / bin_width - base_bin);
    if (bin + 2 > (size_t)nodes) bin = (size_t)nodes - 2;
    else if (bin > 0 && e < bins[bin]) --bin;
    else if (bin + 2 < (size_t)nodes && e > bins[bin + 1]) ++bin;
    return bin;
  }
  float value(float e) const {
    if (e <= bins[0]) return data[0];
    if (e >= bins[nodes - 1]) return data[nodes - 1];
    size_t id = 0;
    if (e < bins[1]) id = 0;
    else if (e >= bins[nodes - 2]) id = (size_t)nodes - 2;
    else if (e < bins[0] || e > bins[1]) id = bin_location(e);
    return interp(bins[id], data[id], bins[id + 1], data[id + 1], e);
  }
};

// GGEMSRangeCuts::ComputePhotonCrossSection (GGEMSRangeCuts.cc:343-403)
float absorption_cross_section(int z, float energy) {
  const float t1keV = 1.0f * unit::keV, t200keV = 200.f * unit::keV, t100MeV = 100.f * unit::MeV;
  const float gZ = static_cast<float>(z);
  const float kZsquare = gZ * gZ;
  const float kZlog = logf(gZ);
  const float kZlogsquare = kZlog * kZlog;
  const float s200keV = (0.2651f - 0.1501f * kZlog + 0.02283f * kZlogsquare) * kZsquare;
  const float tmin = (0.552f + 218.5f / gZ + 557.17f / kZsquare) * unit::MeV;
  const float smin = (0.01239f + 0.005585f * kZlog - 0.000923f * kZlogsquare) * expf(1.5f * kZlog);
  const float cmin = logf(s200keV / smin) / (logf(tmin / t200keV) * logf(tmin / t200keV));
  const float tlow = 0.2f * expf(-7.355f / sqrtf(gZ)) * unit::MeV;
  const float slow = s200keV * expf(0.042f * gZ * logf(t200keV / tlow) * logf(t200keV / tlow));
  const float s1keV = 300.0f * kZsquare;
  const float clow = logf(s1keV / slow) / logf(tlow / t1keV);
  const float chigh = (7.55e-5f - 0.0542e-5f * gZ) * kZsquare * gZ / logf(t100MeV / tmin);
  float xs;
  if (energy < tlow) {
    xs = energy < t1keV ? slow * expf(clow * logf(tlow / t1keV)) : slow * expf(clow * logf(tlow / energy));
  } else if (energy < t200keV) {
    xs = s200keV * expf(0.042f * gZ * logf(t200keV / energy) * logf(t200keV / energy));
  } else if (energy < tmin) {
    xs = smin * expf(cmin * logf(tmin / energy) * logf(tmin / energy));
  } else {
    xs = smin + chigh * logf(energy / tmin);
  }
  return xs * unit::barn;
}
}  // namespace

float photon_energy_cut(const MaterialTables& mt, int mat, float cut, float emin) {
  const float emax = 10.0f * unit::GeV;
  const int nbins = 300;
  const int ne = mt.n_elements[mat], off = mt.element_index[mat];
  // BuildElementsLossTable (:291-322) + BuildAbsorptionLengthTable (:244-264)
  std::vector<LogTable> elem;
  for (int i = 0; i < ne; ++i) {
    elem.emplace_back(emin, emax, nbins);
    for (int j = 0; j < nbins; ++j) elem[i].data[j] = absorption_cross_section(mt.z[off + i], elem[i].bins[j]);
  }
  LogTable range(emin, emax, nbins);
  for (int i = 0; i < nbins; ++i) {
    float sigma = 0.0f;
    for (int j = 0; j < ne; ++j) sigma += mt.atom_density[off + j] * elem[j].data[i];
    range.data[i] = 5.0f / sigma;
  }
  // ConvertLengthToEnergyCut (:534-584)
  const float epsilon = 0.01f;
  float rmax = -1.e10f * unit::mm, t1 = emin, r1 = range.data[0], t2 = emax, kec;
  if (cut < r1) {
    kec = t1;
  } else {
    for (int i = 0; i < nbins; ++i) {
      float t = range.bins[i], r = range.data[i];
      if (r > rmax) rmax = r;
      if (r < cut) {
        t1 = t;
        r1 = r;
      } else if (r > cut) {
        t2 = t;
        break;
      }
    }
    if (cut >= rmax) {
      kec = emax;
    } else {
      float t3 = sqrtf(t1 * t2), r3 = range.value(t3);
      while (fabsf(1.0f - r3 / cut) > epsilon) {
        if (cut <= r3) t2 = t3;
        else t1 = t3;
        t3 = sqrtf(t1 * t2);
        r3 = range.value(t3);
      }
      kec = t3;
    }
  }
  if (kec < emin) kec = emin;
  else if (kec > emax) kec = emax;
  return kec;
}

// ---------------------------------------------------------------------------
// cross sections
// ---------------------------------------------------------------------------
std::vector<float> energy_grid(int nbins, float emin, float emax) {
  // GGEMSCrossSections::Initialize (GGEMSCrossSections.cc:203-207)
  std::vector<float> e(nbins);
  float slope = logf(emax / emin);
  for (int i = 0; i < nbins; ++i)
    e[i] = emin * expf(slope * (static_cast<float>(i) / (static_cast<float>(nbins) - 1.0f))) * unit::MeV;
  return e;
}

// GGEMSComptonScattering::ComputeCrossSectionPerAtom (GGEMSComptonScattering.cc:75-101)
float cs_compton_atom(float energy, int zi) {
  float cs = 0.0f;
  const unsigned char z = static_cast<unsigned char>(zi);
  if (z < 1 || energy < 1e-4f) return cs;
  float p1Z = z * (2.7965e-23f + 1.9756e-27f * z + -3.9178e-29f * z * z);
  float p2Z = z * (-1.8300e-23f + -1.0205e-24f * z + 6.8241e-27f * z * z);
  float p3Z = z * (6.7527e-22f + -7.3913e-24f * z + 6.0480e-27f * z * z);
  float p4Z = z * (-1.9798e-21f + 2.7079e-24f * z + 3.0274e-26f * z * z);
  float T0 = (z < 1.5f) ? 40.0e-3f : 15.0e-3f;
  float d1 = fmaxf(energy, T0) / kElectronMassC2;
  cs = p1Z * logf(1.0f + 2.0f * d1) / d1 +
       (p2Z + p3Z * d1 + p4Z * d1 * d1) / (1.0f + 20.0f * d1 + 230.0f * d1 * d1 + 440.0f * d1 * d1 * d1);
  if (energy < T0) {
    d1 = (T0 + 1.0e-3f) / kElectronMassC2;
    float d2 = p1Z * logf(1.0f + 2.0f * d1) / d1 +
               (p2Z + p3Z * d1 + p4Z * d1 * d1) / (1.0f + 20.0f * d1 + 230.0f * d1 * d1 + 440.0f * d1 * d1 * d1);
    float d3 = (-T0 * (d2 - cs)) / (cs * 1.0e-3f);
    float d4 = (z > 1.5f) ? 0.375f - 0.0556f * logf(static_cast<float>(z)) : 0.15f;
    float d5 = logf(energy / T0);
    cs *= expf(-d5 * (d3 + d4 * d5));
  }
  return cs;
}

// GGEMSPhotoElectricEffect::ComputeCrossSectionPerAtom (GGEMSPhotoElectricEffect.cc:76-98)
float cs_photoelectric_atom(float energy, int z) {
  const PhysicsData& pd = PhysicsData::get();
  static const float* ip = pd.f32("sandia_ionization_potential");
  static const int32_t* nint = pd.i32("sandia_number_of_intervals");
  static const int32_t* cum = pd.i32("sandia_cumulative_intervals");
  static const float* ztoa = pd.f32("sandia_z_to_a_ratio");
  static const float* tab = pd.f32("sandia_table");
  float e_min = std::fmax(ip[z], 10.0f) * unit::eV;
  if (energy < e_min) return 0.0f;
  int start = cum[z - 1];
  int pos = start + nint[z];
  while (energy < static_cast<float>(tab[pos * 5]) * unit::keV) --pos;
  float aover_avo = kAtomicMassUnit * static_cast<float>(z) / ztoa[z];
  double inv = 1.0 / static_cast<double>(energy);
  double inv2 = inv * inv;
  const double a = static_cast<double>(aover_avo);
  return static_cast<float>(inv * static_cast<double>(tab[pos * 5 + 1]) * a * 0.160217648e-22 +
                            inv2 * static_cast<double>(tab[pos * 5 + 2]) * a * 0.160217648e-25 +
                            inv * inv2 * static_cast<double>(tab[pos * 5 + 3]) * a * 0.160217648e-28 +
                            inv2 * inv2 * static_cast<double>(tab[pos * 5 + 4]) * a * 0.160217648e-31);
}

// GGEMSRayleighScattering::ComputeCrossSectionPerAtom (GGEMSRayleighScattering.cc:80-102)
// with LogLogInterpolation (include/GGEMS/maths/GGEMSMathAlgorithms.hh:88-98)
float cs_rayleigh_atom(float energy, int z) {
  if (energy < 250e-6f || energy > 100e3f) return 0.0f;
  const PhysicsData& pd = PhysicsData::get();
  static const int32_t* cum = pd.i32("rayleigh_cs_cumulative_intervals");
  static const int32_t* nint = pd.i32("rayleigh_cs_number_of_intervals");
  static const float* cs = pd.f32("rayleigh_cross_section");
  const int start = cum[z];
  const int stop = start + 2 * (nint[z] - 1);
  int pos = start;
  for (; pos < stop; pos += 2)
    if (cs[pos] >= energy) break;
  if (energy < 1e3f) {
    float x = energy, x0 = cs[pos - 2], y0 = cs[pos - 1], x1 = cs[pos], y1 = cs[pos + 1], v;
    if (x < x0) v = y0;
    else if (x > x1) v = y1;
    else {
      x0 = 1.0f / x0;
      v = powf(10.0f, log10f(y0) + log10f(y1 / y0) * (log10f(x * x0) / log10f(x1 * x0)));
    }
    return 1.0e-22f * v;
  }
  return 1.0e-22f * cs[pos - 1];
}

float cs_atom(int process, float e, int z) {
  switch (process) {
    case kCompton: return cs_compton_atom(e, z);
    case kPhotoelectric: return cs_photoelectric_atom(e, z);
    default: return cs_rayleigh_atom(e, z);
  }
}

// GGEMSEMProcess::BuildCrossSectionTables / ComputeCrossSectionPerMaterial (GGEMSEMProcess.cc:65-143)
CrossSectionTables build_cross_sections(const MaterialTables& mt, const std::vector<int>& processes, int nbins,
                                        float emin, float emax) {
  CrossSectionTables t;
  t.nbins = nbins;
  t.emin = emin;
  t.emax = emax;
  t.processes = processes;
  t.energy_bins = energy_grid(nbins, emin, emax);
  const int nm = (int)mt.names.size();
  t.cs.assign((size_t)3 * nm * nbins, 0.0f);
  t.cs_atom.assign((size_t)3 * 101 * nbins, 0.0f);
  for (int p : processes) {
    for (int m = 0; m < nm; ++m) {
      for (int i = 0; i < nbins; ++i) {
        const float e = t.energy_bins[i];
        float acc = 0.0f;
        const int off = mt.element_index[m];
        for (int k = 0; k < mt.n_elements[m]; ++k) {
          const int z = mt.z[off + k];
          const float a = cs_atom(p, e, z);
          t.cs_atom[((size_t)p * 101 + z) * nbins + i] = a;
          acc += mt.atom_density[off + k] * a;
        }
        t.cs[((size_t)p * nm + m) * nbins + i] = acc;
      }
    }
  }
  return t;
}

namespace {
int bsearch_bins(float key, const float* a, int size) {
  int low = 0, mid = 0, high = size - 1;
  while (high - low > 1) {
    mid = (high + low) / 2;
    if (a[mid] == key) return mid;
    if (a[mid] < key) low = mid;
    else high = mid;
  }
  return low;
}
float lin(float xa, float ya, float xb, float yb, float x) {
  if (xa > xb) return yb;
  if (xa >= x) return ya;
  if (xb <= x) return yb;
  return ya + (x - xa) * (yb - ya) / (xb - xa);
}
}  // namespace

// GGEMSCrossSections::GetPhotonCrossSection (GGEMSCrossSections.cc:263-337)
float photon_cross_section(const CrossSectionTables& t, const MaterialTables& mt, int process, int mat, float e) {
  if (e < t.emin || e > t.emax) raise("GGEMSCrossSections", "energy outside the cross-section table range");
  const int nm = (int)mt.names.size();
  const int b = bsearch_bins(e, t.energy_bins.data(), t.nbins);
  const float* row = &t.cs[((size_t)process * nm + mat) * t.nbins];
  float cs = lin(t.energy_bins[b], row[b], t.energy_bins[b + 1], row[b + 1], e);
  return (cs / mt.density[mat]) / (unit::cm * unit::cm / unit::g);
}

// ---------------------------------------------------------------------------
// attenuation (GGEMSAttenuations.cc:41-231)
// ---------------------------------------------------------------------------
AttenuationTables build_attenuations(const MaterialTables& mt) {
  const PhysicsData& pd = PhysicsData::get();
  const int32_t* nbe = pd.i32("mu_number_of_energy_bins");
  const int32_t* idx_e = pd.i32("mu_index_energy");
  const float* data = pd.f32("mu_data");
  std::vector<float> energies, mus, mu_ens;
  size_t id = 0;
  for (int i = 0; i <= 92; ++i)
    for (int j = 0; j < nbe[i]; ++j) {
      energies.push_back(data[id++]);
      mus.push_back(data[id++]);
      mu_ens.push_back(data[id++]);
    }
  AttenuationTables t;
  const int nb = 220;
  const float emin = 0.001f, emax = 1.0f;
  const float slope = logf(emax / emin);
  for (int i = 0; i < nb; ++i)
    t.energy_bins.push_back(emin * expf(slope * (static_cast<float>(i) / (static_cast<float>(nb) - 1.0f))) * unit::MeV);
  const int nm = (int)mt.names.size();
  t.mu.assign((size_t)nm * nb, 0.0f);
  t.mu_en.assign((size_t)nm * nb, 0.0f);
  for (int m = 0; m < nm; ++m)
    for (int i = 0; i < nb; ++i) {
      const float e = t.energy_bins[i];
      float mu = 0.0f, muen = 0.0f;
      for (int k = 0; k < mt.n_elements[m]; ++k) {
        const int z = mt.z[mt.element_index[m] + k];
        const float frac = mt.mass_fraction[mt.element_index[m] + k];
        const int mie = idx_e[z];
        // the search spans the concatenated energy lists up to element z, as the reference does
        const int ei = bsearch_bins(e, energies.data(), mie + nbe[z]);
        if (ei == mie) {
          mu += mus[ei];
          muen += mu_ens[ei];
        } else {
          mu += frac * lin(energies[ei - 1], mus[ei - 1], energies[ei], mus[ei], e);
          muen += frac * lin(energies[ei - 1], mu_ens[ei - 1], energies[ei], mu_ens[ei], e);
        }
      }
      t.mu[(size_t)m * nb + i] = mu * mt.density[m] / (unit::g / unit::cm3);
      t.mu_en[(size_t)m * nb + i] = muen * mt.density[m] / (unit::g / unit::cm3);
    }
  return t;
}

float attenuation(const AttenuationTables& t, int mat, float e, bool en) {
  const int nb = (int)t.energy_bins.size();
  if (e < 0.001f || e > 1.0f) raise("GGEMSAttenuations", "energy outside [1 keV, 1 MeV]");
  const int b = bsearch_bins(e, t.energy_bins.data(), nb);
  const std::vector<float>& v = en ? t.mu_en : t.mu;
  return lin(t.energy_bins[b], v[(size_t)mat * nb + b], t.energy_bins[b + 1], v[(size_t)mat * nb + b + 1], e);
}

int process_id(const std::string& name) {
  if (name == "Compton") return kCompton;
  if (name == "Photoelectric") return kPhotoelectric;
  if (name == "Rayleigh") return kRayleigh;
  raise("GGEMSCrossSections",
        "Unknown process!!! The available processes in GGEMS are: 'Compton', 'Photoelectric', 'Rayleigh'");
}

const char* process_name(int id) {
  return id == kCompton ? "Compton" : id == kPhotoelectric ? "Photoelectric" : "Rayleigh";
}

}  // namespace gghost
