/*
 * ggo_tables.c -- oracle restatement of the GGEMS host-side table builders.
 * TEST INFRASTRUCTURE ONLY (see ggems_oracle.h).
 *
 * Follows, in the reference's float operation order:
 *   src/materials/GGEMSMaterialsDatabaseManager.cc:97-141  (database parse)
 *   src/io/GGEMSTextReader.cc:43-127                        (line rules)
 *   include/GGEMS/materials/GGEMSMaterialsDatabaseManager.hh:244-255 (atom density)
 *   src/materials/GGEMSMaterials.cc:209-300                 (material tables)
 *   src/physics/GGEMSCrossSections.cc:170-219               (energy grid)
 *   src/physics/GGEMSEMProcess.cc:65-143                    (sigma per material)
 *   src/physics/GGEMSComptonScattering.cc:75-101            (Compton sigma_atom)
 *   src/physics/GGEMSPhotoElectricEffect.cc:76-98           (Sandia sigma_atom)
 *   src/physics/GGEMSRayleighScattering.cc:80-102           (Livermore sigma_atom)
 *   src/physics/GGEMSRangeCuts.cc:129-238,343-403,534-628   (photon cut)
 *   src/physics/GGEMSLogEnergyTable.cc:39-161               (log table)
 *   src/physics/GGEMSAttenuations.cc:41-231                 (mu, mu_en)
 */
#include "ggems_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* units and constants (include/GGEMS/tools/GGEMSSystemOfUnits.hh:54-175,
 * include/GGEMS/global/GGEMSConstants.hh) */
#define U_EV 1.e-6f
#define U_KEV 1.e-3f
#define U_MEV 1.f
#define U_GEV 1.e+3f
#define U_G 6.241509704e+21f
#define U_MG 6.241509704e+18f
#define U_CM3 1.e3f
#define U_BARN 1.e-22f
#define C_AVOGADRO 6.02214179e+23f
#define C_ELECTRON_MASS_C2 0.510998910f
#define C_ATOMIC_MASS_UNIT 0.01036426891f

/* ------------------------------------------------------------------------ */
/* physics data blob                                                        */
/* ------------------------------------------------------------------------ */
static unsigned char* g_blob = NULL;
static uint32_t g_nentries = 0;

int ggo_load_physics_data(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  unsigned char* b = (unsigned char*)malloc((size_t)n);
  if (!b || fread(b, 1, (size_t)n, f) != (size_t)n) { fclose(f); free(b); return -2; }
  fclose(f);
  if (memcmp(b, "GGPD", 4) != 0) { free(b); return -3; }
  free(g_blob);
  g_blob = b;
  memcpy(&g_nentries, b + 8, 4);
  return 0;
}

const void* ggo_data(const char* name, uint64_t* count) {
  if (!g_blob) return NULL;
  for (uint32_t i = 0; i < g_nentries; ++i) {
    const unsigned char* e = g_blob + 12 + (size_t)i * 72;
    if (strncmp((const char*)e, name, 48) == 0) {
      uint64_t cnt, off;
      memcpy(&cnt, e + 56, 8);
      memcpy(&off, e + 64, 8);
      if (count) *count = cnt;
      return g_blob + off;
    }
  }
  return NULL;
}

#define DATA(type, name) ((const type*)ggo_data(name, NULL))

/* ------------------------------------------------------------------------ */
/* materials database                                                       */
/* ------------------------------------------------------------------------ */
typedef struct {
  char name[64];
  float density;
  int n;
  char elem[GGO_MAX_ELEMENTS][32];
  float frac[GGO_MAX_ELEMENTS];
} db_material;

static db_material* g_db = NULL;
static int g_ndb = 0;

/* GGEMSTextReader::RemoveSpace (src/io/GGEMSTextReader.cc:112-127) */
static void remove_space(char* s) {
  char* w = s;
  for (char* r = s; *r; ++r)
    if (*r != '\t' && *r != ' ' && *r != '\n' && *r != '\r' && *r != '\f' && *r != '\v') *w++ = *r;
  *w = 0;
}

static int is_blank(const char* s) {
  for (; *s; ++s)
    if (*s != '\t' && *s != '\n' && *s != ' ') return 0;
  return 1;
}

static const char* first_nonblank(const char* s) {
  while (*s == '\t' || *s == ' ') ++s;
  return *s ? s : NULL;
}

static int read_line(FILE* f, char* buf, int n) {
  if (!fgets(buf, n, f)) return 0;
  size_t l = strlen(buf);
  if (l && buf[l - 1] == '\n') buf[l - 1] = 0;
  return 1;
}

/* GGEMSMaterialReader::ReadMaterialDensity (src/io/GGEMSTextReader.cc:52-77) */
static float read_density(const char* line, int* err) {
  const char* d = strstr(line, "d=");
  const char* p = d ? d : line;
  p += strcspn(p, "0123456789");
  const char* q = p + strspn(p, "0123456789.");
  char num[64];
  size_t ln = (size_t)(q - p) < sizeof(num) - 1 ? (size_t)(q - p) : sizeof(num) - 1;
  memcpy(num, p, ln);
  num[ln] = 0;
  float density = strtof(num, NULL);
  const char* semi = strchr(line, ';');
  char unit[32];
  size_t lu = semi ? (size_t)(semi - q) : strlen(q);
  if (lu > sizeof(unit) - 1) lu = sizeof(unit) - 1;
  memcpy(unit, q, lu);
  unit[lu] = 0;
  if (strcmp(unit, "g/cm3") == 0) density *= U_G / U_CM3;
  else if (strcmp(unit, "mg/cm3") == 0) density *= U_MG / U_CM3;
  else *err = 1;
  return density;
}

int ggo_load_materials(const char* path) {
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  free(g_db);
  g_db = NULL;
  g_ndb = 0;
  int cap = 0;
  char line[1024];
  while (read_line(f, line, sizeof line)) {
    /* SkipComment (src/io/GGEMSTextReader.cc:89-98) */
    for (;;) {
      const char* p = first_nonblank(line);
      if (!p) break;
      if (*p == '#') {
        if (!read_line(f, line, sizeof line)) { line[0] = 0; break; }
      } else break;
    }
    if (is_blank(line)) continue;
    remove_space(line);
    db_material m;
    memset(&m, 0, sizeof m);
    const char* colon = strchr(line, ':');
    size_t ln = colon ? (size_t)(colon - line) : strlen(line);
    if (ln > sizeof(m.name) - 1) ln = sizeof(m.name) - 1;
    memcpy(m.name, line, ln);
    int err = 0;
    m.density = read_density(line, &err);
    if (err) { fclose(f); return -2; }
    const char* np = strstr(line, "n=");
    np = np ? np : line;
    np += strcspn(np, "0123456789");
    m.n = atoi(np);
    if (m.n > GGO_MAX_ELEMENTS) { fclose(f); return -3; }
    for (int i = 0; i < m.n; ++i) {
      if (!read_line(f, line, sizeof line)) { fclose(f); return -4; }
      remove_space(line);
      const char* nm = strstr(line, "name=");
      nm = nm ? nm + 5 : line;
      const char* semi = strchr(line, ';');
      size_t le = semi && semi > nm ? (size_t)(semi - nm) : strlen(nm);
      if (le > 31) le = 31;
      memcpy(m.elem[i], nm, le);
      m.elem[i][le] = 0;
      const char* fp = strstr(line, "f=");
      fp = fp ? fp : line;
      fp += strcspn(fp, "0123456789");
      m.frac[i] = strtof(fp, NULL);
    }
    int dup = 0;
    for (int k = 0; k < g_ndb; ++k)
      if (strcmp(g_db[k].name, m.name) == 0) dup = 1; /* std::map::insert keeps the first */
    if (dup) continue;
    if (g_ndb == cap) {
      cap = cap ? cap * 2 : 256;
      g_db = (db_material*)realloc(g_db, (size_t)cap * sizeof(db_material));
    }
    g_db[g_ndb++] = m;
  }
  fclose(f);
  return g_ndb;
}

static const db_material* find_material(const char* name) {
  for (int k = 0; k < g_ndb; ++k)
    if (strcmp(g_db[k].name, name) == 0) return &g_db[k];
  return NULL;
}

static int find_element(const char* name, int* z, float* molar_mass) {
  uint64_t n = 0;
  const char* names = (const char*)ggo_data("element_names", &n);
  const int32_t* zs = DATA(int32_t, "element_z");
  const float* ms = DATA(float, "element_molar_mass");
  if (!names || !zs || !ms) return -1;
  int idx = 0;
  const char* p = names;
  while (*p) {
    const char* e = strchr(p, '\n');
    size_t l = e ? (size_t)(e - p) : strlen(p);
    if (strlen(name) == l && strncmp(name, p, l) == 0) {
      *z = zs[idx];
      /* molar_mass_M_ = element_M * (g / mol) (GGEMSMaterialsDatabaseManager.cc:463) */
      *molar_mass = ms[idx] * (U_G / 1.0f);
      return 0;
    }
    if (!e) break;
    p = e + 1;
    ++idx;
  }
  return -1;
}

int ggo_material_info(const char* name, float* density, int* n_elements, uint8_t* z_out,
                      float* fraction_out) {
  const db_material* m = find_material(name);
  if (!m) return -1;
  *density = m->density;
  *n_elements = m->n;
  for (int i = 0; i < m->n; ++i) {
    int z;
    float mm;
    if (find_element(m->elem[i], &z, &mm)) return -2;
    z_out[i] = (uint8_t)z;
    fraction_out[i] = m->frac[i];
  }
  return 0;
}

int ggo_build_material_tables(const char* const* names, int n, float cut_distance,
                              float cut_emin, ggo_material_tables* out) {
  memset(out, 0, sizeof *out);
  out->n_materials = n;
  int idx = 0;
  for (int i = 0; i < n; ++i) {
    const db_material* m = find_material(names[i]);
    if (!m) return -1 - i;
    out->n_elements[i] = m->n;
    out->density[i] = m->density;
    out->elem_index[i] = idx;
    for (int j = 0; j < m->n; ++j) {
      int z;
      float mm;
      if (find_element(m->elem[j], &z, &mm)) return -1000 - i;
      out->z[idx + j] = (uint8_t)z;
      out->mass_fraction[idx + j] = m->frac[j];
      /* GetAtomicNumberDensity (.hh:244-255): FP64 then narrowed to FP32 */
      double v = (double)C_AVOGADRO / (double)mm * (double)m->density * (double)m->frac[j];
      out->atom_density[idx + j] = (float)v;
    }
    idx += m->n;
  }
  if (cut_distance > 0.0f)
    for (int i = 0; i < n; ++i) out->photon_cut[i] = ggo_photon_cut(out, i, cut_distance, cut_emin);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* cross sections                                                           */
/* ------------------------------------------------------------------------ */
/* GGEMSCrossSections::Initialize energy grid (GGEMSCrossSections.cc:203-207) */
void ggo_energy_bins(int nbins, float emin, float emax, float* out) {
  float slope = logf(emax / emin);
  for (int i = 0; i < nbins; ++i)
    out[i] = emin * expf(slope * ((float)i / ((float)nbins - 1.0f))) * U_MEV;
}

/* GGEMSComptonScattering::ComputeCrossSectionPerAtom (GGEMSComptonScattering.cc:75-101) */
float ggo_cs_compton_atom(float energy, int zi) {
  float cs = 0.0f;
  unsigned char z = (unsigned char)zi;
  if (z < 1 || energy < 1e-4f) return cs;
  float p1Z = z * (2.7965e-23f + 1.9756e-27f * z + -3.9178e-29f * z * z);
  float p2Z = z * (-1.8300e-23f + -1.0205e-24f * z + 6.8241e-27f * z * z);
  float p3Z = z * (6.7527e-22f + -7.3913e-24f * z + 6.0480e-27f * z * z);
  float p4Z = z * (-1.9798e-21f + 2.7079e-24f * z + 3.0274e-26f * z * z);
  float T0 = (z < 1.5f) ? 40.0e-3f : 15.0e-3f;
  float d1, d2, d3, d4, d5;
  d1 = fmaxf(energy, T0) / C_ELECTRON_MASS_C2;
  cs = p1Z * logf(1.0f + 2.0f * d1) / d1 +
       (p2Z + p3Z * d1 + p4Z * d1 * d1) / (1.0f + 20.0f * d1 + 230.0f * d1 * d1 + 440.0f * d1 * d1 * d1);
  if (energy < T0) {
    d1 = (T0 + 1.0e-3f) / C_ELECTRON_MASS_C2;
    d2 = p1Z * logf(1.0f + 2.0f * d1) / d1 +
         (p2Z + p3Z * d1 + p4Z * d1 * d1) / (1.0f + 20.0f * d1 + 230.0f * d1 * d1 + 440.0f * d1 * d1 * d1);
    d3 = (-T0 * (d2 - cs)) / (cs * 1.0e-3f);
    d4 = (z > 1.5f) ? 0.375f - 0.0556f * logf((float)z) : 0.15f;
    d5 = logf(energy / T0);
    cs *= expf(-d5 * (d3 + d4 * d5));
  }
  return cs;
}

/* GGEMSPhotoElectricEffect::ComputeCrossSectionPerAtom (GGEMSPhotoElectricEffect.cc:76-98) */
float ggo_cs_photoelectric_atom(float energy, int z) {
  const float* ip = DATA(float, "sandia_ionization_potential");
  const int32_t* nint = DATA(int32_t, "sandia_number_of_intervals");
  const int32_t* cum = DATA(int32_t, "sandia_cumulative_intervals");
  const float* ztoa = DATA(float, "sandia_z_to_a_ratio");
  const float* tab = DATA(float, "sandia_table");
  float e_min = fmaxf(ip[z], 10.0f) * U_EV;
  if (energy < e_min) return 0.0f;
  int start = cum[z - 1];
  int stop = start + nint[z];
  int pos = stop;
  while (energy < (float)tab[pos * 5 + 0] * U_KEV) --pos;
  float aover_avo = C_ATOMIC_MASS_UNIT * (float)z / ztoa[z];
  double energy_inv = 1.0 / (double)energy;
  double energy_inv2 = energy_inv * energy_inv;
  return (float)(energy_inv * (double)tab[pos * 5 + 1] * (double)aover_avo * 0.160217648e-22 +
                 energy_inv2 * (double)tab[pos * 5 + 2] * (double)aover_avo * 0.160217648e-25 +
                 energy_inv * energy_inv2 * (double)tab[pos * 5 + 3] * (double)aover_avo * 0.160217648e-28 +
                 energy_inv2 * energy_inv2 * (double)tab[pos * 5 + 4] * (double)aover_avo * 0.160217648e-31);
}

/* LogLogInterpolation (include/GGEMS/maths/GGEMSMathAlgorithms.hh:88-98, host branch) */
static float loglog_interp(float x, float x0, float y0, float x1, float y1) {
  if (x < x0) return y0;
  if (x > x1) return y1;
  x0 = 1.0f / x0;
  return powf(10.0f, log10f(y0) + log10f(y1 / y0) * (log10f(x * x0) / log10f(x1 * x0)));
}

/* GGEMSRayleighScattering::ComputeCrossSectionPerAtom (GGEMSRayleighScattering.cc:80-102) */
float ggo_cs_rayleigh_atom(float energy, int z) {
  if (energy < 250e-6f || energy > 100e3f) return 0.0f;
  const int32_t* cum = DATA(int32_t, "rayleigh_cs_cumulative_intervals");
  const int32_t* nint = DATA(int32_t, "rayleigh_cs_number_of_intervals");
  const float* cs = DATA(float, "rayleigh_cross_section");
  int start = cum[z];
  int stop = start + 2 * (nint[z] - 1);
  int pos = start;
  for (; pos < stop; pos += 2)
    if (cs[pos] >= energy) break;
  if (energy < 1e3f)
    return 1.0e-22f * loglog_interp(energy, cs[pos - 2], cs[pos - 1], cs[pos], cs[pos + 1]);
  return 1.0e-22f * cs[pos - 1];
}

static float cs_atom(int process, float e, int z) {
  if (process == 0) return ggo_cs_compton_atom(e, z);
  if (process == 1) return ggo_cs_photoelectric_atom(e, z);
  return ggo_cs_rayleigh_atom(e, z);
}

/* GGEMSEMProcess::BuildCrossSectionTables + ComputeCrossSectionPerMaterial
 * (GGEMSEMProcess.cc:65-143) */
void ggo_cs_per_material(int process, const ggo_material_tables* mt, int nbins,
                         const float* ebins, float* cs_out, float* cs_atom_out) {
  for (int m = 0; m < mt->n_materials; ++m) {
    for (int i = 0; i < nbins; ++i) {
      float e = ebins[i];
      float acc = 0.0f;
      int off = mt->elem_index[m];
      for (int k = 0; k < mt->n_elements[m]; ++k) {
        int z = mt->z[off + k];
        float a = cs_atom(process, e, z);
        if (cs_atom_out) cs_atom_out[z * nbins + i] = a;
        acc += mt->atom_density[off + k] * a;
      }
      cs_out[m * nbins + i] = acc;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* photon range cut                                                         */
/* ------------------------------------------------------------------------ */
#define LT_BINS 300
typedef struct {
  float bin_width, base_bin;
  int nodes;
  float bins[LT_BINS + 1];
  float data[LT_BINS + 1];
} log_table;

/* GGEMSLogEnergyTable::GGEMSLogEnergyTable (GGEMSLogEnergyTable.cc:39-64) */
static void lt_init(log_table* t, float lo, float hi, int nbins) {
  t->bin_width = logf(hi / lo) / (float)nbins;
  t->base_bin = logf(lo) / t->bin_width;
  t->nodes = nbins + 1;
  t->bins[0] = lo;
  t->data[0] = 0.0f;
  for (int i = 1; i < t->nodes - 1; ++i) {
    t->bins[i] = expf((t->base_bin + (float)i) * t->bin_width);
    t->data[i] = 0.0f;
  }
  t->bins[t->nodes - 1] = hi;
  t->data[t->nodes - 1] = 0.0f;
}

/* LinearInterpolation (include/GGEMS/maths/GGEMSMathAlgorithms.hh:67-77) */
static float lin_interp(float xa, float ya, float xb, float yb, float x) {
  if (xa > xb) return yb;
  if (xa >= x) return ya;
  if (xb <= x) return yb;
  return ya + (x - xa) * (yb - ya) / (xb - xa);
}

/* GGEMSLogEnergyTable::FindBinLocation / FindBin / GetLossTableValue (:88-161) */
static size_t lt_find_bin_location(const log_table* t, float e) {
  size_t bin = (size_t)(logf(e) / t->bin_width - t->base_bin);
  if (bin + 2 > (size_t)t->nodes) bin = (size_t)t->nodes - 2;
  else if (bin > 0 && e < t->bins[bin]) --bin;
  else if (bin + 2 < (size_t)t->nodes && e > t->bins[bin + 1]) ++bin;
  return bin;
}

static float lt_value(const log_table* t, float e) {
  size_t idx = 0;
  if (e <= t->bins[0]) return t->data[0];
  if (e >= t->bins[t->nodes - 1]) return t->data[t->nodes - 1];
  if (e < t->bins[1]) idx = 0;
  else if (e >= t->bins[t->nodes - 2]) idx = (size_t)t->nodes - 2;
  else if (e < t->bins[0] || e > t->bins[1]) idx = lt_find_bin_location(t, e);
  return lin_interp(t->bins[idx], t->data[idx], t->bins[idx + 1], t->data[idx + 1], e);
}

/* GGEMSRangeCuts::ComputePhotonCrossSection (GGEMSRangeCuts.cc:343-403) */
static float photon_abs_cs(int z, float energy) {
  float t1keV = 1.0f * U_KEV;
  float t200keV = 200.f * U_KEV;
  float t100MeV = 100.f * U_MEV;
  float gZ = (float)z;
  float kZsquare = gZ * gZ;
  float kZlog = logf(gZ);
  float kZlogsquare = kZlog * kZlog;
  float s200keV = (0.2651f - 0.1501f * kZlog + 0.02283f * kZlogsquare) * kZsquare;
  float tmin = (0.552f + 218.5f / gZ + 557.17f / kZsquare) * U_MEV;
  float smin = (0.01239f + 0.005585f * kZlog - 0.000923f * kZlogsquare) * expf(1.5f * kZlog);
  float cmin = logf(s200keV / smin) / (logf(tmin / t200keV) * logf(tmin / t200keV));
  float tlow = 0.2f * expf(-7.355f / sqrtf(gZ)) * U_MEV;
  float slow = s200keV * expf(0.042f * gZ * logf(t200keV / tlow) * logf(t200keV / tlow));
  float s1keV = 300.0f * kZsquare;
  float clow = logf(s1keV / slow) / logf(tlow / t1keV);
  float chigh = (7.55e-5f - 0.0542e-5f * gZ) * kZsquare * gZ / logf(t100MeV / tmin);
  float xs;
  if (energy < tlow) {
    if (energy < t1keV) xs = slow * expf(clow * logf(tlow / t1keV));
    else xs = slow * expf(clow * logf(tlow / energy));
  } else if (energy < t200keV) {
    xs = s200keV * expf(0.042f * gZ * logf(t200keV / energy) * logf(t200keV / energy));
  } else if (energy < tmin) {
    xs = smin * expf(cmin * logf(tmin / energy) * logf(tmin / energy));
  } else {
    xs = smin + chigh * logf(energy / tmin);
  }
  return xs * U_BARN;
}

/* GGEMSRangeCuts::ConvertToEnergy, gamma branch (GGEMSRangeCuts.cc:129-238)
 * with BuildElementsLossTable / BuildAbsorptionLengthTable / ConvertLengthToEnergyCut */
float ggo_photon_cut(const ggo_material_tables* mt, int mat, float cut, float emin) {
  const float emax = 10.0f * U_GEV;
  int ne = mt->n_elements[mat];
  int off = mt->elem_index[mat];
  static __thread log_table elem[GGO_MAX_ELEMENTS];
  for (int i = 0; i < ne; ++i) {
    lt_init(&elem[i], emin, emax, LT_BINS);
    for (int j = 0; j < LT_BINS; ++j) elem[i].data[j] = photon_abs_cs(mt->z[off + i], elem[i].bins[j]);
  }
  log_table range;
  lt_init(&range, emin, emax, LT_BINS);
  for (int i = 0; i < LT_BINS; ++i) {
    float sigma = 0.0f;
    for (int j = 0; j < ne; ++j) sigma += mt->atom_density[off + j] * elem[j].data[i];
    range.data[i] = 5.0f / sigma;
  }
  /* ConvertLengthToEnergyCut (GGEMSRangeCuts.cc:534-584) */
  float epsilon = 0.01f;
  float rmax = -1.e10f;
  float t1 = emin;
  float r1 = range.data[0];
  float t2 = emax;
  float kec;
  if (cut < r1) {
    kec = t1;
  } else {
    for (int i = 0; i < LT_BINS; ++i) {
      float t = range.bins[i];
      float r = range.data[i];
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
      float t3 = sqrtf(t1 * t2);
      float r3 = lt_value(&range, t3);
      while (fabsf(1.0f - r3 / cut) > epsilon) {
        if (cut <= r3) t2 = t3;
        else t1 = t3;
        t3 = sqrtf(t1 * t2);
        r3 = lt_value(&range, t3);
      }
      kec = t3;
    }
  }
  if (kec < emin) kec = emin;
  else if (kec > emax) kec = emax;
  return kec;
}

/* ------------------------------------------------------------------------ */
/* attenuation (GGEMSAttenuations.cc:41-231)                                */
/* ------------------------------------------------------------------------ */
static int bsearch_f(float key, const float* a, int size) {
  int low = 0, mid = 0, high = size - 1;
  while (high - low > 1) {
    mid = (high + low) / 2;
    if (a[mid] == key) return mid;
    if (a[mid] < key) low = mid;
    else high = mid;
  }
  return low;
}

/* The reference looks each energy up with BinarySearch over the
 * concatenated per-element energy lists from index 0 up to the end of
 * element Z's list (GGEMSAttenuations.cc:200-201).  The lists restart at
 * 1 keV for every element, so for Z > 1 the bisection runs over an unsorted
 * array and can land in another element's list; the resulting table is
 * what GGEMS returns, and the oracle reproduces it (element_search = 0).
 * element_search = 1 searches element Z's own list instead: the physical
 * mixture rule, for the pins against the NIST tables. */
static void attenuation_table(const ggo_material_tables* mt, int mat, int element_search, float* ebins, float* mu,
                              float* mu_en);

void ggo_attenuation(const ggo_material_tables* mt, int mat, float* ebins, float* mu, float* mu_en) {
  attenuation_table(mt, mat, 0, ebins, mu, mu_en);
}

void ggo_attenuation_element_search(const ggo_material_tables* mt, int mat, float* ebins, float* mu, float* mu_en) {
  attenuation_table(mt, mat, 1, ebins, mu, mu_en);
}

static void attenuation_table(const ggo_material_tables* mt, int mat, int element_search, float* ebins, float* mu,
                              float* mu_en) {
  const int32_t* nbe = DATA(int32_t, "mu_number_of_energy_bins");
  const int32_t* idx_e = DATA(int32_t, "mu_index_energy");
  const float* data = DATA(float, "mu_data");
  static float energies[3467], mus[3467], mu_ens[3467];
  int it = 0, id = 0;
  for (int i = 0; i <= 92; ++i)
    for (int j = 0; j < nbe[i]; ++j) {
      energies[it] = data[id++];
      mus[it] = data[id++];
      mu_ens[it] = data[id++];
      ++it;
    }
  const float emin = 0.001f, emax = 1.0f;
  float slope = logf(emax / emin);
  for (int i = 0; i < 220; ++i) ebins[i] = emin * expf(slope * ((float)i / ((float)220 - 1.0f))) * U_MEV;
  for (int i = 0; i < 220; ++i) {
    float e = ebins[i];
    float m = 0.0f, men = 0.0f;
    for (int k = 0; k < mt->n_elements[mat]; ++k) {
      int z = mt->z[mt->elem_index[mat] + k];
      float frac = mt->mass_fraction[mt->elem_index[mat] + k];
      int mie = idx_e[z];
      int ei = element_search ? mie + bsearch_f(e, energies + mie, nbe[z]) : bsearch_f(e, energies, mie + nbe[z]);
      if (ei == mie) {
        m += mus[ei];
        men += mu_ens[ei];
      } else {
        m += frac * lin_interp(energies[ei - 1], mus[ei - 1], energies[ei], mus[ei], e);
        men += frac * lin_interp(energies[ei - 1], mu_ens[ei - 1], energies[ei], mu_ens[ei], e);
      }
    }
    mu[i] = m * mt->density[mat] / (U_G / U_CM3);
    mu_en[i] = men * mt->density[mat] / (U_G / U_CM3);
  }
}
