/*
 * ggems_oracle.h -- CPU oracle for the GGEMS photon-transport hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used by tests/ as the checker and by bench.py's cpu_baseline leg
 * as the timed CPU baseline.  The product (ggems_amd) never links, loads or
 * calls it.
 *
 * Every function cites the reference file:line it follows.  Arithmetic is
 * written in the reference's operation order and built with
 * -ffp-contract=off so that host-side tables are bit-exact restatements.
 */
#ifndef GGEMS_ORACLE_H
#define GGEMS_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- physics data blob (ggems_amd/data/_gen/ggems_physics.bin) ---------- */
int ggo_load_physics_data(const char* blob_path);
/* returns pointer to array `name` and its element count, or NULL */
const void* ggo_data(const char* name, uint64_t* count);

/* ---- materials database (src/materials/GGEMSMaterialsDatabaseManager.cc) - */
int ggo_load_materials(const char* materials_txt);
int ggo_material_info(const char* name, float* density, int* n_elements,
                      uint8_t* z_out, float* fraction_out);

/* ---- tables for one navigator ------------------------------------------ */
#define GGO_MAX_MATERIALS 255
#define GGO_MAX_ELEMENTS 32
#define GGO_MAX_BINS 2048

typedef struct {
  int n_materials;
  int n_elements[GGO_MAX_MATERIALS];
  int elem_index[GGO_MAX_MATERIALS];
  float density[GGO_MAX_MATERIALS];
  float photon_cut[GGO_MAX_MATERIALS];
  uint8_t z[GGO_MAX_MATERIALS * GGO_MAX_ELEMENTS];
  float atom_density[GGO_MAX_MATERIALS * GGO_MAX_ELEMENTS];
  float mass_fraction[GGO_MAX_MATERIALS * GGO_MAX_ELEMENTS];
} ggo_material_tables;

/* builds material tables (GGEMSMaterials::BuildMaterialTables) and, when
 * cut_distance > 0, the photon energy cuts (GGEMSRangeCuts::ConvertToEnergy,
 * gamma branch) with table minimum `cut_emin` */
int ggo_build_material_tables(const char* const* names, int n, float cut_distance,
                              float cut_emin, ggo_material_tables* out);

void ggo_energy_bins(int nbins, float emin, float emax, float* out);
float ggo_cs_compton_atom(float energy, int z);
float ggo_cs_photoelectric_atom(float energy, int z);
float ggo_cs_rayleigh_atom(float energy, int z);
/* process: 0 Compton, 1 photoelectric, 2 Rayleigh.  cs_out[m*nbins + i];
 * cs_atom_out[z*nbins + i] (101*nbins, may be NULL) */
void ggo_cs_per_material(int process, const ggo_material_tables* mt, int nbins,
                         const float* energy_bins, float* cs_out, float* cs_atom_out);
float ggo_photon_cut(const ggo_material_tables* mt, int mat, float distance, float emin);
/* GGEMSAttenuations::Initialize mu / mu_en (220 bins, 1 keV..1 MeV) */
void ggo_attenuation(const ggo_material_tables* mt, int mat, float* ebins220,
                     float* mu220, float* mu_en220);
/* the same table with each element's energy looked up in that element's own
 * list (the physical mixture rule; the reference searches the concatenated
 * lists of all elements up to Z, GGEMSAttenuations.cc:200-201) */
void ggo_attenuation_element_search(const ggo_material_tables* mt, int mat, float* ebins220,
                                    float* mu220, float* mu_en220);

/* ---- RNG (include/GGEMS/randoms/GGEMSKissEngine.hh) --------------------- */
void ggo_mt19937_stream(uint32_t seed, uint64_t n, uint32_t* out);
void ggo_seed_slots(uint32_t seed, uint64_t n_slots, uint32_t* states /* 5*n_slots, SoA */);
float ggo_kiss_uniform(uint32_t* s /* 5 words */);

/* ---- geometry (include/GGEMS/geometries/GGEMSRayTracing.hh) ------------ */
typedef struct {
  float m[4][4];          /* GGfloat44 rows m0_..m3_ */
  float bmin[3], bmax[3]; /* local borders */
} ggo_obb;

float ggo_distance_to_aabb(const float* p, const float* d, const float* bmin, const float* bmax);

/* ---- scene -------------------------------------------------------------- */
enum { GGO_SOLID_VOXEL = 0, GGO_SOLID_BOX = 1, GGO_SOLID_MESH = 2 };

/* meshed solid (GGEMSMeshedSolid): triangles in global coordinates and the
 * GGEMSOctree over them, with each node's triangle list as indices */
typedef struct {
  int n_triangles;
  const float* tri;        /* [n_triangles][9]: p0, p1, p2 */
  int n_nodes;
  const float* node_box;   /* [n_nodes][6]: centre xyz, half width xyz */
  const int* node_first;   /* [n_nodes]: first triangle of the node's list, -1 if none */
  const int* tri_next;     /* [n_triangles]: next triangle in the same list, -1 at the end */
} ggo_mesh;

typedef struct {
  int type;
  ggo_obb obb;
  int mat_offset;        /* global material index of this navigator's material 0 */
  /* voxel solid */
  float voxel_size[3];
  int nvox[3];
  const uint8_t* labels;
  /* dosimetry (voxel or meshed solid) */
  int dosimetry;
  float dosel_size[3], inv_dosel[3], dose_bmin[3], dose_bmax[3];
  int ndosel[3];
  double* edep;
  double* edep2;  /* may be NULL */
  int32_t* hit;   /* may be NULL */
  int32_t* photon_tracking; /* may be NULL */
  /* box solid (CT module) */
  int nelem[3];
  float box_size[3];
  float threshold;
  int32_t* histogram;
  int32_t* scatter;  /* may be NULL */
  /* meshed solid */
  ggo_mesh mesh;
} ggo_solid;

typedef struct {
  int nbins;
  const float* energy_bins;
  int n_processes;
  int process_id[3];            /* activation order */
  const float* cs[3];           /* [process][gmat*nbins + bin], NULL if inactive */
  const float* cs_atom_rayleigh;/* [z*nbins + bin] (101*nbins) */
  /* global material tables */
  const int* n_elements;
  const int* elem_index;
  const uint8_t* z;
  const float* atom_density;
  const float* photon_cut;
  /* 0: the reference's sampling, three competing exponentials per voxel
   * step; 1: one exponential number of mean free paths per flight, spent
   * voxel by voxel (opt-in, no counterpart in the reference; see
   * track_voxel) */
  int tau_mode;
} ggo_physics;

typedef struct {
  float m[4][4];
  float aperture;
  float focal[3];
  int n_energy;           /* number_of_energy_bins_ (spectrum lines) */
  const float* energy;    /* n_energy+1 entries */
  const float* cdf;       /* n_energy+1 entries */
  int is_interp;
} ggo_source;

typedef struct {
  uint64_t primaries;
  uint64_t voxel_steps;      /* S: voxel-loop iterations in voxel solids */
  uint64_t interactions;     /* I: discrete processes in voxel solids */
  uint64_t detector_steps;
  uint64_t detector_interactions; /* D: discrete processes in box solids */
  uint64_t hops;
  uint64_t mesh_steps;
  uint64_t mesh_interactions;
} ggo_stats;

/* Transports n_particles primaries with the reference's batch/slot structure
 * (src/global/GGEMS.cc:303-363, src/sources/GGEMSSource.cc:371-426):
 * slot s owns JKISS state s and processes photon s of every batch.
 * threads <= 0 uses all host cores. */
int ggo_run(const ggo_physics* ph, const ggo_solid* solids, int n_solids,
            const ggo_source* src, uint64_t n_particles, uint32_t seed,
            int threads, ggo_stats* stats);

/* Same, restricted to slots [slot_begin, slot_end) (bounded CPU-baseline
 * samples).  Returns the number of primaries transported. */
uint64_t ggo_run_slots(const ggo_physics* ph, const ggo_solid* solids, int n_solids,
                       const ggo_source* src, uint64_t n_particles, uint32_t seed,
                       uint64_t slot_begin, uint64_t slot_end, int threads, ggo_stats* stats);

/* CT acquisition as consecutive runs: projection p places the source at
 * src_mats[p] and box solid j (rank among the box solids) at
 * box_mats[p][j] (row-major 4x4) and transports n_particles[p] primaries in
 * the reference's batch layout; slot s runs its primaries of projection 0,
 * 1, ... in order, continuing its JKISS state as consecutive GGEMS::Run
 * calls would (src/global/GGEMS.cc:303-363).  The histogram/scatter of
 * every box solid points at n_projections consecutive images.  phantoms
 * (NULL, or one entry per projection): projection p runs in a world of the
 * box solids and only solid phantoms[p] of the others (-1: every solid). */
int ggo_run_acquisition(const ggo_physics* ph, const ggo_solid* solids, int n_solids, const ggo_source* src,
                        int n_projections, const float* src_mats, const float* box_mats,
                        const uint64_t* n_particles, const int32_t* phantoms, uint32_t seed, int threads,
                        ggo_stats* stats);

/* number of batches / slots for n particles (reference batch split) */
void ggo_batch_layout(uint64_t n_particles, uint64_t* n_batches, uint64_t* n_slots);

/* ---- single-function entry points for tests ---------------------------- */
/* voxel indices visited by a ray through a voxel solid with no interaction
 * (pure TRANSPORTATION steps, TrackThroughGGEMSVoxelizedSolid.cl:133-310);
 * position/direction are global.  Returns the number of voxels visited. */
int ggo_voxel_walk(const ggo_solid* solid, const float* p, const float* d,
                   int32_t* out_index, int max_out);
/* one Compton / Rayleigh interaction on (E, dir); returns new E, writes dir */
float ggo_sample_compton(float e, float* dir, uint32_t* rng);
float ggo_sample_rayleigh(const ggo_physics* ph, int gmat, float e, float* dir,
                          uint32_t* rng, int* status);
/* one source primary: pos/dir/E */
void ggo_sample_primary(const ggo_source* src, uint32_t* rng, float* pos, float* dir, float* e);

/* GGEMSMeshedSolid::BuildOctree over triangles already moved to global
 * coordinates (src/geometries/GGEMSMeshedSolid.cc:296-458,
 * src/geometries/GGEMSOctree.cc:40-316, src/io/GGEMSSTLReader.cc:111-226):
 * fills node_box [(8^depth-1)/7][6], node_first, tri_next and the octree
 * box (the OBB borders).  Returns the node count or -1. */
int ggo_mesh_build(const float* tri, int n_triangles, int depth, float* node_box, int* node_first,
                   int* tri_next, float* obb_min, float* obb_max);
/* GetMeshIOPositions (TrackThroughGGEMSMeshSolid.cl:253-314): number of
 * ray/mesh intersections and the (up to) 16 nearest points, sorted */
int ggo_mesh_hits(const ggo_mesh* mesh, const float* p, const float* d, float* points /* [16][3] */);

/* compute_dose_ggems_voxelized_solid (ComputeDoseGGEMSVoxelizedSolid.cl:53-121) */
void ggo_compute_dose(const ggo_solid* phantom, const float* density_by_gmat,
                      float scale, int water_reference, float min_density,
                      float* dose, float* uncertainty);

#ifdef __cplusplus
}
#endif
#endif
