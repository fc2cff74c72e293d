/*
 * ggems_engine.h -- thin C ABI of the MI355X photon-transport engine.
 *
 * This is the inner drop-in boundary: the entry points GGEMS's host code
 * would call in place of its OpenCL launch sites.  Plain C types only
 * (pointers, sizes, POD descriptors); every call returns 0 on success or a
 * negative code, with a message from ggh_last_error().  Nothing throws.
 *
 * Reference launch sites each entry point replaces:
 *   ggh_set_physics       GGEMSCrossSections::Initialize + GGEMSMaterials::BuildMaterialTables
 *                         uploads (src/physics/GGEMSCrossSections.cc:170-219,
 *                         src/materials/GGEMSMaterials.cc:209-300)
 *   ggh_add_voxel_solid   GGEMSVoxelizedSolid::LoadVolumeImage / UpdateTransformationMatrix
 *                         (src/geometries/GGEMSVoxelizedSolid.cc:217-317)
 *   ggh_add_box_solid     GGEMSSolidBox constructor + histogram allocation
 *                         (src/geometries/GGEMSSolidBox.cc:40-112)
 *   ggh_set_dosimetry     GGEMSDosimetryCalculator::Initialize
 *                         (src/navigators/GGEMSDosimetryCalculator.cc:435-531)
 *   ggh_set_xray_source   GGEMSXRaySource::Initialize / FillEnergy
 *                         (src/sources/GGEMSXRaySource.cc:241-256, src/sources/GGEMSSource.cc:265-316)
 *   ggh_seed              GGEMSPseudoRandomGenerator::InitializeSeeds
 *                         (src/randoms/GGEMSPseudoRandomGenerator.cc:140-167)
 *   ggh_run               GGEMS::RunOnDevice batch/hop loop: get_primaries, particle_solid_distance_*,
 *                         project_to_*, track_through_*, is_alive
 *                         (src/global/GGEMS.cc:303-363, src/navigators/GGEMSNavigator.cc:320-530,
 *                          src/sources/GGEMSXRaySource.cc:100-148, src/physics/GGEMSParticles.cc:142-193)
 *   ggh_compute_dose      GGEMSDosimetryCalculator::ComputeDose
 *                         (src/navigators/GGEMSDosimetryCalculator.cc:346-430)
 *   ggh_get_histogram     GGEMSSystem::SaveResults readback (src/navigators/GGEMSSystem.cc:165-256)
 *   ggh_get_dose_tallies  GGEMSDosimetryCalculator::Save* readback
 *                         (src/navigators/GGEMSDosimetryCalculator.cc:536-809)
 */
#ifndef GGEMS_ENGINE_H
#define GGEMS_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GGH_MAX_SLOTS 1048576   /* particle slots per device (MAXIMUM_PARTICLES) */
#define GGH_MAX_SOLIDS 256
#define GGH_MAX_HOPS 100        /* navigator hops per primary (src/global/GGEMS.cc:330) */

typedef struct ggh_context ggh_context;

/* Oriented box: local->global 4x4 matrix, row major (GGfloat44 m0_..m3_),
 * and the local-frame borders. */
typedef struct {
  float matrix[16];
  float border_min[3];
  float border_max[3];
} ggh_obb;

/* Physics tables for every material of every navigator, concatenated.
 * A navigator's material i is global material (material_offset + i). */
typedef struct {
  int32_t nbins;                 /* energy bins of the cross-section tables */
  int32_t n_materials;           /* global material count */
  int32_t n_processes;           /* activated photon processes */
  int32_t process_id[3];         /* activation order: 0 Compton, 1 photoelectric, 2 Rayleigh */
  const float* energy_bins;      /* [nbins] MeV */
  const float* cross_sections;   /* [3][n_materials][nbins] mm^-1 (rows of inactive processes ignored) */
  const float* cs_atom_rayleigh; /* [101][nbins] mm^2 (Rayleigh per atom, element selection) */
  int32_t total_elements;
  const int32_t* n_elements;     /* [n_materials] */
  const int32_t* element_index;  /* [n_materials] offset into the element arrays */
  const uint8_t* atomic_number;  /* [total_elements] */
  const float* atom_density;     /* [total_elements] atoms/mm^3 */
  const float* photon_cut;       /* [n_materials] MeV */
  const float* density;          /* [n_materials] internal density units */
  const float* rayleigh_pp;      /* [9][101] form-factor fit parameters PP0..PP8 */
  float rayleigh_factor;         /* 0.5*(hc)^-2 constant of the angular sampler */
} ggh_physics_desc;

typedef struct {
  ggh_obb obb;
  int32_t material_offset;
  int32_t n_voxels[3];
  float voxel_size[3];
  const uint8_t* labels;         /* host pointer, n_voxels[0]*n_voxels[1]*n_voxels[2] bytes */
} ggh_voxel_solid_desc;

typedef struct {
  ggh_obb obb;
  int32_t material_offset;
  int32_t n_elements[3];         /* virtual detection elements */
  float box_size[3];
  float threshold;               /* MeV */
  int32_t store_scatter;
} ggh_box_solid_desc;

/* Meshed solid (GGEMSMeshedSolid): a closed triangle mesh of one material,
 * in global coordinates (the OBB matrix is the identity and its borders are
 * the octree box, src/geometries/GGEMSMeshedSolid.cc:296-458). */
typedef struct {
  ggh_obb obb;
  int32_t material_offset;
  int32_t n_triangles;
  const float* triangles;        /* host pointer, [n_triangles][3 points][3] */
} ggh_mesh_solid_desc;

typedef struct {
  int32_t solid;                 /* voxel or meshed solid (ggh_add_voxel_solid, ggh_add_mesh_solid) */
  float dosel_size[3];
  int32_t n_dosels[3];
  float border_min[3];
  float border_max[3];
  int32_t edep_squared;          /* keep sum of edep^2 */
  int32_t hit;                   /* keep hit counts */
  int32_t photon_tracking;       /* keep photon fluence counts */
} ggh_dosimetry_desc;

typedef struct {
  float matrix[16];              /* source local->global */
  float aperture;                /* rad */
  float focal_spot[3];           /* mm */
  int32_t n_energy;              /* spectrum lines; energy/cdf hold n_energy+1 entries */
  const float* energy;
  const float* cdf;
  int32_t is_interp;
} ggh_xray_source_desc;

/* run statistics (cumulative since the last ggh_reset_tallies) */
typedef struct {
  uint64_t primaries;
  uint64_t voxel_steps;
  uint64_t interactions;         /* discrete processes inside voxel solids */
  uint64_t detector_steps;
  uint64_t detector_interactions;/* discrete processes inside box solids */
  uint64_t hops;
  uint64_t mesh_steps;           /* iterations of the meshed-solid transport loop */
  uint64_t mesh_interactions;    /* discrete processes inside meshed solids */
} ggh_stats;

const char* ggh_last_error(void);
int ggh_device_count(int* count);

int ggh_context_create(int device, ggh_context** ctx);
int ggh_context_destroy(ggh_context* ctx);

int ggh_set_physics(ggh_context* ctx, const ggh_physics_desc* desc);
/* both return the new solid's index (>= 0) or a negative error; solids are
 * visited in the order they are added, as GGEMS registers them */
int ggh_add_voxel_solid(ggh_context* ctx, const ggh_voxel_solid_desc* desc);
int ggh_add_box_solid(ggh_context* ctx, const ggh_box_solid_desc* desc);
/* meshed solid (GGEMSMeshedSolid::Initialize + BuildOctree, src/navigators/GGEMSMeshedPhantom.cc:80-155) */
int ggh_add_mesh_solid(ggh_context* ctx, const ggh_mesh_solid_desc* desc);
int ggh_set_dosimetry(ggh_context* ctx, const ggh_dosimetry_desc* desc);
int ggh_set_xray_source(ggh_context* ctx, const ggh_xray_source_desc* desc);

/* seeds every particle slot from std::mt19937(seed); seed 0 draws from /dev/urandom */
int ggh_seed(ggh_context* ctx, uint32_t seed);
/* transports n primaries (batches of GGH_MAX_SLOTS, slot s owns RNG stream s);
 * tallies accumulate across calls */
int ggh_run(ggh_context* ctx, uint64_t n_particles);
int ggh_synchronize(ggh_context* ctx);
/* milliseconds of device time of the last ggh_run transport kernel */
int ggh_last_kernel_ms(ggh_context* ctx, float* ms);
int ggh_reset_tallies(ggh_context* ctx);
int ggh_get_stats(ggh_context* ctx, ggh_stats* out);

/* launch shape of ggh_run (no counterpart in GGEMS; for tuning and tests).
 * Zero fields keep the defaults. */
typedef struct {
  int32_t round_steps;     /* voxel steps a wave makes at least between two regroupings of its
                              block's photons (1..1024; default 20 with the reference sampling, 16
                              with tau sampling) */
  int32_t cold_passes;     /* navigation passes per round for the photons that are not inside a
                              phantom (1..16; default 2 with the reference sampling, 1 with tau
                              sampling) */
  int32_t interact_batch;  /* inside the voxel loop, resolve pending Compton/Rayleigh/photoelectric
                              interactions once this many lanes wait for one (default 16; 64 leaves
                              them to the navigation passes) */
  int32_t blocks_per_cu;   /* cap on persistent blocks per compute unit (default: as many as fit) */
  int32_t general_kernel;  /* nonzero: the general kernel variant even where the specialised one applies */
} ggh_launch_desc;
int ggh_set_launch(ggh_context* ctx, const ggh_launch_desc* desc);

/* How a voxel step samples its interaction (no counterpart in GGEMS).
 * GGH_SAMPLING_REFERENCE (the default) draws the three competing
 * exponentials of GetPhotonNextInteraction in every voxel step
 * (include/GGEMS/navigators/GGEMSPhotonNavigator.hh:57-93), so every history
 * is GGEMS's.  GGH_SAMPLING_TAU draws one exponential number of mean free
 * paths per flight (on entering a voxel solid and after each interaction)
 * and spends sigma_tot * db of it per crossed voxel; the process is chosen
 * by one uniform in proportion to its left-bin cross section.  The voxel
 * walk, tolerances, interactions and tallies are unchanged and the
 * distributions per voxel are the same; the random streams, and so the
 * individual histories, differ from GGEMS's.  Boxes and meshes keep the
 * reference sampling. */
#define GGH_SAMPLING_REFERENCE 0
#define GGH_SAMPLING_TAU 1
int ggh_set_sampling(ggh_context* ctx, int32_t mode);
/* grid and kernel variant of the last ggh_run: 0 specialised (three processes,
 * one voxel solid), 1 general, 2 three processes over several solids */
int ggh_last_grid(ggh_context* ctx, int32_t* blocks, int32_t* block_size, int32_t* variant);

int ggh_get_histogram(ggh_context* ctx, int solid, int32_t* counts, int32_t* scatter);

/* ---- CT acquisition ------------------------------------------------------
 * A gantry of n_projections positions transported in ONE persistent launch.
 * The reference runs one projection per GGEMS::Run after placing the source
 * and the CT system with SetRotation (src/navigators/GGEMSCTSystem.cc:
 * 279-288, src/global/GGEMS.cc:303-363); here projection p places the
 * source at source_matrices[p] and box solid j (registration order among
 * the box solids) at box_matrices[p][j], and transports n_particles[p]
 * primaries laid out in batches as ggh_run lays them out.  Slot s runs its
 * primaries of projection 0, then 1, ..., so its random stream continues
 * exactly as over consecutive ggh_run calls: the images equal those of
 * n_projections ggh_run calls with the same placements.  Each box solid
 * scores n_projections images; phantom dose tallies sum over projections.
 * Several phantoms share one launch through `phantoms`: projection p then
 * sees the box solids and only the one other solid phantoms[p] (-1: every
 * solid), as if the other phantoms had not been registered for that run. */
#define GGH_MAX_PROJECTIONS 65536
typedef struct {
  int32_t n_projections;
  const float* source_matrices;  /* [n_projections][16] row-major local->global */
  const float* box_matrices;     /* [n_projections][n_box_solids][16] */
  const uint64_t* n_particles;   /* [n_projections] */
  const int32_t* phantoms;       /* [n_projections] solid index or -1; NULL: every solid in every projection */
} ggh_acquisition_desc;
/* uploads placements and zeroes the images (tallies of earlier acquisitions are dropped) */
int ggh_set_acquisition(ggh_context* ctx, const ggh_acquisition_desc* desc);
/* one launch of the set acquisition; images and dose tallies accumulate */
int ggh_run_acquisition(ggh_context* ctx);
/* image `projection` of box solid `solid` (counts and optionally scatter) */
int ggh_get_projection(ggh_context* ctx, int solid, int projection, int32_t* counts, int32_t* scatter);
/* the set acquisition captured as a hipGraph (slot-queue reset + transport
 * launch); replays continue the random streams like further
 * ggh_run_acquisition calls.  ggh_set_acquisition drops the graph. */
int ggh_capture_acquisition(ggh_context* ctx);
int ggh_replay_acquisition(ggh_context* ctx);

/* Ray-traced primary transmission (DRR) onto the pixels of box solid
 * `solid`: image[row][col] = sum_b w_b exp(-sum_m sigma_tot(m, b) L_m), the
 * probability that a source photon aimed at the pixel centre crosses the
 * voxel solids without interacting.  L_m is the exact path length in
 * material m, sigma_tot the left-bin total cross section the transport uses
 * for its distances (include/GGEMS/navigators/GGEMSPhotonNavigator.hh:57-93),
 * w_b the probability of energy bin b under the source's sampling
 * (src/kernels/GetPrimariesGGEMSXRaySource.cl:103-125).  projection < 0:
 * the set-up placement; otherwise that projection of the set acquisition.
 * No counterpart in GGEMS: it feeds the scatter surrogate. */
int ggh_drr(ggh_context* ctx, int solid, int projection, float* image);

/* ggh_drr without allocating: no counterpart in GGEMS.  For callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops).  `image` is a device buffer of the box solid's rows x cols floats
 * on the context's device; the two kernels (spectrum weights, ray tracer)
 * are queued on `stream` (a hipStream_t; null: the context's stream) and the
 * call returns without waiting.  The weights are staged in one array of
 * 8,448 bytes per GPU, shared by every call on that GPU: calls that overlap
 * on one GPU race.  Limits: at most 1,024 energy bins in the physics tables,
 * and the solid table must already be on the device (after a ggh_run,
 * ggh_run_acquisition or ggh_drr since the last solid was added); otherwise
 * the call fails and allocates nothing.  Same image as ggh_drr, bit for bit. */
int ggh_drr_noalloc(ggh_context* ctx, int solid, int projection, float* image, void* stream);

/* Launch autotuner: times ggh_run(n_trial) over a grid of launch shapes
 * (blocks per compute unit, round steps, and with the reference sampling the
 * interaction batch) and keeps the fastest.  Call after ggh_seed and before any run: it restores the seeded
 * random streams and zeroes every tally and counter afterwards. */
int ggh_autotune(ggh_context* ctx, uint64_t n_trial, ggh_launch_desc* best);
int ggh_get_dose_tallies(ggh_context* ctx, double* edep, double* edep_squared, int32_t* hit,
                         int32_t* photon_tracking);
int ggh_compute_dose(ggh_context* ctx, float scale_factor, int32_t water_reference, float minimum_density,
                     float* dose, float* uncertainty);

/* ---- voxelized volume painter ----------------------------------------------
 * GGEMSVolumeCreatorManager's label volume (src/geometries/
 * GGEMSVolumeCreatorManager.cc:170-232) and its box / tube / sphere painters
 * (src/kernels/DrawGGEMS{Box,Tube,Sphere}.cl), on the device. */
typedef struct ggh_volume ggh_volume;
enum { GGH_MET_CHAR = 0, GGH_MET_UCHAR, GGH_MET_SHORT, GGH_MET_USHORT, GGH_MET_INT, GGH_MET_UINT, GGH_MET_FLOAT };
enum { GGH_SHAPE_BOX = 0, GGH_SHAPE_TUBE = 1, GGH_SHAPE_SPHERE = 2 };
typedef struct {
  int32_t shape;
  float position[3];  /* mm, from the volume centre */
  float label;        /* converted to the element type */
  float size[3];      /* mm.  box: height (y), width (x), depth (z); tube: height (z),
                         radius x, radius y; sphere: radius */
} ggh_volume_shape;
const char* ggh_volume_last_error(void);
/* zero-filled volume of dims[0]*dims[1]*dims[2] elements of data_type (x fastest) */
int ggh_volume_create(int device, const int32_t dims[3], const float element_size[3], int32_t data_type,
                      ggh_volume** out);
/* paints every voxel whose centre lies in the shape (later shapes overwrite earlier ones) */
int ggh_volume_draw(ggh_volume* volume, const ggh_volume_shape* shape);
/* copies the volume to host memory (bytes must be its exact size) */
int ggh_volume_read(ggh_volume* volume, void* host, size_t bytes);
int ggh_volume_destroy(ggh_volume* volume);

/* ---- scatter-surrogate U-Net: 3x3 convolutions on the matrix cores ---------
 * No counterpart in GGEMS (the surrogate is new).  BF16 NHWC activations,
 * BF16 weights [cout][9][cin] (tap = 3 (dy + 1) + (dx + 1)), FP32
 * accumulation; zero padding 1.  Device pointers; `stream` is a hipStream_t
 * (null: the default stream).  W % 16 == 0, N H W % 32 == 0, cin % 8 == 0,
 * cout % 16 == 0. */
const char* ggh_unet_last_error(void);
/* y = act(conv(x, w) + bias); act = ReLU when relu != 0; bias may be null */
int ggh_conv3x3_forward(const void* x, const void* w, const float* bias, void* y, int32_t n, int32_t h, int32_t w_px,
                        int32_t cin, int32_t cout, int32_t relu, void* stream);
/* dw[cout][9 cin] += sum over pixels of dy (x) shifted x, FP32 (cout <= 128);
 * the pixels are split over `splits` groups of blocks that add atomically */
int ggh_conv3x3_weight_grad(const void* x, const void* dy, float* dw, int32_t n, int32_t h, int32_t w_px, int32_t cin,
                            int32_t cout, int32_t splits, void* stream);

/* a layer's BF16 operands from its FP32 weights w [cout][9][cin]: w_bf16
 * [cout][9][cin] (round to nearest even) and, for the data gradient, wt_bf16
 * [cin][9][cout] with the taps reversed (tap 8 - t).  The same pass sets the
 * layer's FP32 gradient accumulators dw_zero [cout][9][cin] and db_zero
 * [cout] to zero; either may be null. */
int ggh_conv3x3_weights_bf16(const float* w, void* w_bf16, void* wt_bf16, float* dw_zero, float* db_zero,
                             int32_t cout, int32_t cin, void* stream);

/* backward of a layer's ReLU and bias: dy_out = dy where y > 0 (everywhere
 * when relu == 0), BF16; db[c] += sum over the pixels of dy_out, FP32.
 * dy is BF16, or FP32 when dy_is_f32; all NHWC with c one of 8, 16, 32, 64, 128 */
int ggh_relu_bias_backward(const void* dy, int32_t dy_is_f32, const void* y, int32_t relu, void* dy_out, float* db,
                           int64_t n_pixels, int32_t c, void* stream);

/* 2 x 2 max pooling, NHWC BF16 x [n][h][w][c] -> y [n][h/2][w/2][c]; arg
 * (one byte per output element) keeps which input of the window (row-major,
 * 0..3) held the first maximum.  h, w even; c % 8 == 0. */
int ggh_pool2x2_forward(const void* x, void* y, uint8_t* arg, int32_t n, int32_t h, int32_t w_px, int32_t c,
                        void* stream);
/* dx [n][h][w][c] = dy at the input arg names, 0 at the other three */
int ggh_pool2x2_backward(const void* dy, const uint8_t* arg, void* dx, int32_t n, int32_t h, int32_t w_px, int32_t c,
                         void* stream);
/* the decoder's input: out [n][h][w][cz + cs] = (z [n][h/2][w/2][cz] repeated
 * over each 2 x 2 block, s [n][h][w][cs]); cz, cs % 8 == 0 */
int ggh_upcat_forward(const void* z, const void* s, void* out, int32_t n, int32_t h, int32_t w_px, int32_t cz,
                      int32_t cs, void* stream);
/* its backward: dz = dout's first cz channels summed over each 2 x 2 block
 * (FP32, rounded once), ds = dout's last cs channels */
int ggh_upcat_backward(const void* dout, void* dz, void* ds, int32_t n, int32_t h, int32_t w_px, int32_t cz,
                       int32_t cs, void* stream);

/* ---- test entry points -------------------------------------------------- */
/* voxel indices visited by rays crossing voxel solid `solid` with no
 * interaction; pos/dir are global.  out_index[r*max_steps + k], out_count[r]. */
int ggh_debug_voxel_walk(ggh_context* ctx, int solid, int n_rays, const float* pos, const float* dir,
                         int max_steps, int32_t* out_index, int32_t* out_count);
/* ray/mesh intersections of meshed solid `solid` (the transport's
 * GetMeshIOPositions): count[r] and the two nearest points nearest[r][2][3] */
int ggh_debug_mesh_hits(ggh_context* ctx, int solid, int n_rays, const float* pos, const float* dir,
                        int32_t* count, float* nearest);
/* the first primary of each of the first n slots (consumes their RNG) */
int ggh_debug_primaries(ggh_context* ctx, int n, float* pos, float* dir, float* energy);
/* n draws from slot 0..n_slots-1 streams: out[s*n + k] */
int ggh_debug_uniforms(ggh_context* ctx, int n_slots, int n, float* out);
/* -log(u) of the voxel step (fast[i]) and -logf(u) (ref[i]) on the device for RNG sums */
int ggh_debug_neg_log(ggh_context* ctx, int n, const uint32_t* sums, float* fast, float* ref);
/* host builds of the engine's arithmetic (exact-equivalence checks against the oracle) */
void ggh_host_uniform_from_sum(const uint32_t* sums, float* out, size_t n);
/* a[i] / b[i] by the engine's division-free sequence */
void ggh_host_div_rn(const float* a, const float* b, float* out, size_t n);
/* the engine's energy-bin lookup (BinarySearch on the energy grid,
 * include/GGEMS/maths/GGEMSMathAlgorithms.hh:46-60) for e[0..n-1] */
void ggh_host_energy_bin(const float* e, size_t n, const float* ebins, int32_t nbins, int32_t* out);
/* the engine's slab test: p, d [n][3], box [n][6] = {min xyz, max xyz} */
void ggh_host_distance_to_aabb(const float* p, const float* d, const float* box, float* out, size_t n);

#ifdef __cplusplus
}
#endif
#endif
