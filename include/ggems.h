/*
 * ggems.h -- GGEMS-compatible C API of libggems.so (the outer drop-in boundary).
 *
 * Same names, argument types and meaning as the functions GGEMS exports
 * with extern "C" and that its ctypes bindings load
 * (python_module/ggems_lib.py:31-36).  Each block cites the reference header
 * whose declarations it reproduces.  Handles are opaque pointers.
 *
 * Errors: GGEMS prints and throws across the C boundary, which ends the
 * process.  Here each call catches the error, prints it, and records it for
 * get_last_error_ggems(); GGEMS_ABORT_ON_ERROR=1 restores the abort.
 */
#ifndef GGEMS_C_API_H
#define GGEMS_C_API_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#else
typedef struct GGEMS GGEMS;
typedef struct GGEMSOpenCLManager GGEMSOpenCLManager;
typedef struct GGEMSMaterials GGEMSMaterials;
typedef struct GGEMSMaterialsDatabaseManager GGEMSMaterialsDatabaseManager;
typedef struct GGEMSRangeCutsManager GGEMSRangeCutsManager;
typedef struct GGEMSCrossSections GGEMSCrossSections;
typedef struct GGEMSProcessesManager GGEMSProcessesManager;
typedef struct GGEMSAttenuations GGEMSAttenuations;
typedef struct GGEMSSourceManager GGEMSSourceManager;
typedef struct GGEMSXRaySource GGEMSXRaySource;
typedef struct GGEMSProfilerManager GGEMSProfilerManager;
typedef struct GGEMSRAMManager GGEMSRAMManager;
typedef struct GGEMSVoxelizedPhantom GGEMSVoxelizedPhantom;
typedef struct GGEMSCTSystem GGEMSCTSystem;
typedef struct GGEMSMeshedPhantom GGEMSMeshedPhantom;
typedef struct GGEMSDosimetryCalculator GGEMSDosimetryCalculator;
typedef struct GGEMSVolumeCreatorManager GGEMSVolumeCreatorManager;
typedef struct GGEMSBox GGEMSBox;
typedef struct GGEMSTube GGEMSTube;
typedef struct GGEMSSphere GGEMSSphere;
#endif

/* ---- include/GGEMS/global/GGEMS.hh:231-334 -------------------------------- */
GGEMS* create_ggems(void);
void delete_ggems(GGEMS* ggems);
void initialize_ggems(GGEMS* ggems, uint32_t seed);
void set_opencl_verbose_ggems(GGEMS* ggems, bool is_opencl_verbose);
void set_material_database_verbose_ggems(GGEMS* ggems, bool is_material_database_verbose);
void set_source_ggems(GGEMS* ggems, bool is_source_verbose);
void set_navigator_ggems(GGEMS* ggems, bool is_navigator_verbose);
void set_memory_ram_ggems(GGEMS* ggems, bool is_memory_ram_verbose);
void set_process_ggems(GGEMS* ggems, bool is_process_verbose);
void set_range_cuts_ggems(GGEMS* ggems, bool is_range_cuts_verbose);
void set_random_ggems(GGEMS* ggems, bool is_random_verbose);
void set_profiling_ggems(GGEMS* ggems, bool is_profiling_verbose);
void set_tracking_ggems(GGEMS* ggems, bool is_tracking_verbose, int particle_id_tracking);
void run_ggems(GGEMS* ggems);

/* ---- include/GGEMS/global/GGEMSOpenCLManager.hh:794-833 (HIP device selection) ---- */
GGEMSOpenCLManager* get_instance_ggems_opencl_manager(void);
void print_infos_opencl_manager(GGEMSOpenCLManager* opencl_manager);
void set_device_index_ggems_opencl_manager(GGEMSOpenCLManager* opencl_manager, size_t device_id);
void set_device_to_activate_opencl_manager(GGEMSOpenCLManager* opencl_manager, char const* device_type,
                                           char const* device_vendor);
void clean_opencl_manager(GGEMSOpenCLManager* opencl_manager);
void set_device_balancing_opencl_manager(GGEMSOpenCLManager* opencl_manager, char const* device_balancing);

/* ---- include/GGEMS/materials/GGEMSMaterials.hh:219-278 --------------------- */
GGEMSMaterials* create_ggems_materials(void);
void add_material_ggems_materials(GGEMSMaterials* materials, char const* material_name);
void initialize_ggems_materials(GGEMSMaterials* materials);
void print_material_properties_ggems_materials(GGEMSMaterials* materials);
float get_density_ggems_materials(GGEMSMaterials* materials, char const* material_name);
float get_energy_cut_ggems_materials(GGEMSMaterials* materials, char const* material_name, char const* particle_type,
                                     float distance, char const* unit);
float get_atomic_number_density_ggems_materials(GGEMSMaterials* materials, char const* material_name);
void clean_ggems_materials(GGEMSMaterials* materials);

/* ---- include/GGEMS/materials/GGEMSMaterialsDatabaseManager.hh:316-338 ------- */
GGEMSMaterialsDatabaseManager* get_instance_materials_manager(void);
void set_materials_ggems_materials_manager(GGEMSMaterialsDatabaseManager* ggems_materials_manager, char const* filename);
void print_available_chemical_elements_ggems_materials_manager(GGEMSMaterialsDatabaseManager* ggems_materials_manager);
void print_available_materials_ggems_materials_manager(GGEMSMaterialsDatabaseManager* ggems_materials_manager);

/* ---- include/GGEMS/physics/GGEMSRangeCutsManager.hh:127-138 ----------------- */
GGEMSRangeCutsManager* get_instance_range_cuts_manager(void);
void set_cut_range_cuts_manager(GGEMSRangeCutsManager* range_cut_manager, char const* phantom_name,
                                char const* particle_name, float value, char const* unit);

/* ---- include/GGEMS/physics/GGEMSCrossSections.hh:173-209 -------------------- */
GGEMSCrossSections* create_ggems_cross_sections(GGEMSMaterials* materials);
void add_process_ggems_cross_sections(GGEMSCrossSections* cross_sections, char const* process_name,
                                      char const* particle_name, bool is_secondary);
void initialize_ggems_cross_sections(GGEMSCrossSections* cross_sections);
float get_cs_ggems_cross_sections(GGEMSCrossSections* cross_sections, char const* process_name,
                                  char const* material_name, float energy, char const* unit);
void clean_ggems_cross_sections(GGEMSCrossSections* cross_sections);

/* ---- include/GGEMS/physics/GGEMSProcessesManager.hh:191-249 ----------------- */
GGEMSProcessesManager* get_instance_processes_manager(void);
void add_process_processes_manager(GGEMSProcessesManager* processes_manager, char const* process_name,
                                   char const* particle_name, char const* phantom_name);
void set_cross_section_table_number_of_bins_processes_manager(GGEMSProcessesManager* processes_manager,
                                                              size_t number_of_bins);
void set_cross_section_table_minimum_energy_processes_manager(GGEMSProcessesManager* processes_manager, float energy,
                                                              char const* unit);
void set_cross_section_table_maximum_energy_processes_manager(GGEMSProcessesManager* processes_manager, float energy,
                                                              char const* unit);
void print_infos_processes_manager(GGEMSProcessesManager* processes_manager);
void print_available_processes_manager(GGEMSProcessesManager* processes_manager);
void print_tables_processes_manager(GGEMSProcessesManager* processes_manager, bool is_processes_print_tables);

/* ---- include/GGEMS/physics/GGEMSAttenuations.hh:157-193 --------------------- */
GGEMSAttenuations* create_ggems_attenuations(GGEMSMaterials* materials, GGEMSCrossSections* cross_sections);
void initialize_ggems_attenuations(GGEMSAttenuations* attenuations);
float get_mu_ggems_attenuations(GGEMSAttenuations* attenuations, char const* material_name, float energy,
                                char const* unit);
float get_mu_en_ggems_attenuations(GGEMSAttenuations* attenuations, char const* material_name, float energy,
                                   char const* unit);
void clean_ggems_attenuations(GGEMSAttenuations* attenuations);

/* ---- include/GGEMS/sources/GGEMSSourceManager.hh:228-250 -------------------- */
GGEMSSourceManager* get_instance_ggems_source_manager(void);
void initialize_source_manager(GGEMSSourceManager* source_manager, uint32_t seed);
void print_infos_source_manager(GGEMSSourceManager* source_manager);
void clean_source_manager(GGEMSSourceManager* source_manager);

/* ---- include/GGEMS/sources/GGEMSXRaySource.hh:145-230 ----------------------- */
GGEMSXRaySource* create_ggems_xray_source(char const* source_name);
void set_position_ggems_xray_source(GGEMSXRaySource* xray_source, float pos_x, float pos_y, float pos_z,
                                    char const* unit);
void set_number_of_particles_xray_source(GGEMSXRaySource* xray_source, size_t number_of_particles);
void set_source_particle_type_ggems_xray_source(GGEMSXRaySource* xray_source, char const* particle_name);
void set_beam_aperture_ggems_xray_source(GGEMSXRaySource* xray_source, float beam_aperture, char const* unit);
void set_focal_spot_size_ggems_xray_source(GGEMSXRaySource* xray_source, float width, float height, float depth,
                                           char const* unit);
void set_rotation_ggems_xray_source(GGEMSXRaySource* xray_source, float rx, float ry, float rz, char const* unit);
void set_monoenergy_ggems_xray_source(GGEMSXRaySource* xray_source, float monoenergy, char const* unit);
void set_polyenergy_ggems_xray_source(GGEMSXRaySource* xray_source, char const* energy_spectrum);
void set_energy_peak_ggems_xray_source(GGEMSXRaySource* xray_source, float energy, char const* unit, float intensity);

/* ---- include/GGEMS/tools/GGEMSPrint.hh:266, GGEMSProfilerManager.hh:138-145, GGEMSRAMManager.hh:174-181 ---- */
void set_ggems_verbose(int verbosity);
GGEMSProfilerManager* get_instance_profiler_manager(void);
void print_summary_profiler_manager(GGEMSProfilerManager* profiler_manager);
GGEMSRAMManager* get_instance_ggems_ram_manager(void);
void print_infos_ram_manager(GGEMSRAMManager* ram_manager);

/* ---- include/GGEMS/navigators/GGEMSVoxelizedPhantom.hh:119-187 -------------- */
GGEMSVoxelizedPhantom* create_ggems_voxelized_phantom(char const* voxelized_phantom_name);
void set_phantom_file_ggems_voxelized_phantom(GGEMSVoxelizedPhantom* voxelized_phantom, char const* phantom_filename,
                                              char const* range_data_filename);
void set_position_ggems_voxelized_phantom(GGEMSVoxelizedPhantom* voxelized_phantom, float position_x,
                                          float position_y, float position_z, char const* unit);
void set_rotation_ggems_voxelized_phantom(GGEMSVoxelizedPhantom* voxelized_phantom, float rx, float ry, float rz,
                                          char const* unit);
void set_visible_ggems_voxelized_phantom(GGEMSVoxelizedPhantom* voxelized_phantom, bool flag);
void set_material_visible_ggems_voxelized_phantom(GGEMSVoxelizedPhantom* voxelized_phantom, char const* material_name,
                                                  bool flag);
void set_material_color_ggems_voxelized_phantom(GGEMSVoxelizedPhantom* voxelized_phantom, char const* material_name,
                                                unsigned char red, unsigned char green, unsigned char blue);
void set_material_color_name_ggems_voxelized_phantom(GGEMSVoxelizedPhantom* voxelized_phantom,
                                                     char const* material_name, char const* color_name);

/* ---- include/GGEMS/navigators/GGEMSMeshedPhantom.hh:125-199 ----------------- */
GGEMSMeshedPhantom* create_ggems_meshed_phantom(char const* meshed_phantom_name);
void set_phantom_file_ggems_meshed_phantom(GGEMSMeshedPhantom* meshed_phantom, char const* phantom_filename);
void set_position_ggems_meshed_phantom(GGEMSMeshedPhantom* meshed_phantom, float position_x, float position_y,
                                       float position_z, char const* unit);
void set_rotation_ggems_meshed_phantom(GGEMSMeshedPhantom* meshed_phantom, float rx, float ry, float rz,
                                       char const* unit);
void set_visible_ggems_meshed_phantom(GGEMSMeshedPhantom* meshed_phantom, bool flag);
void set_material_name_ggems_meshed_phantom(GGEMSMeshedPhantom* meshed_phantom, char const* material_name);
void set_material_color_ggems_meshed_phantom(GGEMSMeshedPhantom* meshed_phantom, char const* material_name,
                                             unsigned char red, unsigned char green, unsigned char blue);
void set_material_color_name_ggems_meshed_phantom(GGEMSMeshedPhantom* meshed_phantom, char const* material_name,
                                                  char const* color_name);
void set_mesh_octree_depth_ggems_meshed_phantom(GGEMSMeshedPhantom* meshed_phantom, int depth);

/* ---- include/GGEMS/navigators/GGEMSCTSystem.hh:141-280 ---------------------- */
GGEMSCTSystem* create_ggems_ct_system(char const* ct_system_name);
void set_number_of_modules_ggems_ct_system(GGEMSCTSystem* ct_system, size_t module_x, size_t module_y);
void set_ct_system_type_ggems_ct_system(GGEMSCTSystem* ct_system, char const* ct_system_type);
void set_number_of_detection_elements_ggems_ct_system(GGEMSCTSystem* ct_system, size_t n_detection_element_x,
                                                      size_t n_detection_element_y, size_t n_detection_element_z);
void set_size_of_detection_elements_ggems_ct_system(GGEMSCTSystem* ct_system, float size_of_detection_element_x,
                                                    float size_of_detection_element_y,
                                                    float size_of_detection_element_z, char const* unit);
void set_material_name_ggems_ct_system(GGEMSCTSystem* ct_system, char const* material_name);
void set_source_isocenter_distance_ggems_ct_system(GGEMSCTSystem* ct_system, float source_isocenter_distance,
                                                   char const* unit);
void set_source_detector_distance_ggems_ct_system(GGEMSCTSystem* ct_system, float source_detector_distance,
                                                  char const* unit);
void set_rotation_ggems_ct_system(GGEMSCTSystem* ct_system, float rx, float ry, float rz, char const* unit);
void set_threshold_ggems_ct_system(GGEMSCTSystem* ct_system, float threshold, char const* unit);
void set_save_ggems_ct_system(GGEMSCTSystem* ct_system, char const* basename);
void store_scatter_ggems_ct_system(GGEMSCTSystem* ct_system, bool is_scatter);
void set_visible_ggems_ct_system(GGEMSCTSystem* ct_system, bool flag);
void set_material_color_ggems_ct_system(GGEMSCTSystem* ct_system, char const* material_name, unsigned char red,
                                        unsigned char green, unsigned char blue);
void set_material_color_name_ggems_ct_system(GGEMSCTSystem* ct_system, char const* material_name,
                                             char const* color_name);
void set_global_system_position_ggems_ct_system(GGEMSCTSystem* ct_system, float global_system_position_x,
                                                float global_system_position_y, float global_system_position_z,
                                                char const* unit);

/* ---- include/GGEMS/navigators/GGEMSDosimetryCalculator.hh:318-425 ----------- */
GGEMSDosimetryCalculator* create_ggems_dosimetry_calculator(void);
void delete_dosimetry_calculator(GGEMSDosimetryCalculator* dose_calculator);
void scale_factor_dosimetry_calculator(GGEMSDosimetryCalculator* dose_calculator, float scale_factor);
void water_reference_dosimetry_calculator(GGEMSDosimetryCalculator* dose_calculator, bool is_activated);
void minimum_density_dosimetry_calculator(GGEMSDosimetryCalculator* dose_calculator, float minimum_density,
                                          char const* unit);
void set_dosel_size_dosimetry_calculator(GGEMSDosimetryCalculator* dose_calculator, float dose_x, float dose_y,
                                         float dose_z, char const* unit);
void set_dose_output_dosimetry_calculator(GGEMSDosimetryCalculator* dose_calculator, char const* dose_output_filename);
void dose_photon_tracking_dosimetry_calculator(GGEMSDosimetryCalculator* dose_calculator, bool is_activated);
void dose_edep_dosimetry_calculator(GGEMSDosimetryCalculator* dose_calculator, bool is_activated);
void dose_hit_dosimetry_calculator(GGEMSDosimetryCalculator* dose_calculator, bool is_activated);
void dose_edep_squared_dosimetry_calculator(GGEMSDosimetryCalculator* dose_calculator, bool is_activated);
void dose_uncertainty_dosimetry_calculator(GGEMSDosimetryCalculator* dose_calculator, bool is_activated);
void dose_tle_navigator(GGEMSDosimetryCalculator* dose_calculator, bool is_activated);
void attach_to_navigator_dosimetry_calculator(GGEMSDosimetryCalculator* dose_calculator, char const* navigator);

/* ---- additions with no counterpart in GGEMS (tests, benchmarks, tools) ----- */
/* message of the most recent failed call on this thread not yet read ("" if none);
 * reading it clears it, so call it after each call whose success matters */
/* ---- volume creator (src/geometries/GGEMSVolumeCreatorManager.cc:262-384,
 * GGEMSBox.cc:171-231, GGEMSTube.cc:171-231, GGEMSSphere.cc:167-227) ---- */
GGEMSVolumeCreatorManager* get_instance_volume_creator_manager(void);
void set_volume_dimension_volume_creator_manager(GGEMSVolumeCreatorManager* volume_creator_manager,
                                                 size_t volume_width, size_t volume_height, size_t volume_depth);
void set_element_sizes_volume_creator_manager(GGEMSVolumeCreatorManager* volume_creator_manager, float voxel_width,
                                              float voxel_height, float voxel_depth, char const* unit);
void set_output_image_filename_volume_creator_manager(GGEMSVolumeCreatorManager* volume_creator_manager,
                                                      char const* output_image_filename);
void set_output_range_to_material_filename_volume_creator_manager(GGEMSVolumeCreatorManager* volume_creator_manager,
                                                                  char const* output_range_to_material_filename);
void initialize_volume_creator_manager(GGEMSVolumeCreatorManager* volume_creator_manager);
void write_volume_creator_manager(GGEMSVolumeCreatorManager* volume_creator_manager);
void set_material_volume_creator_manager(GGEMSVolumeCreatorManager* volume_creator_manager, char const* material);
void set_data_type_volume_creator_manager(GGEMSVolumeCreatorManager* volume_creator_manager, char const* data_type);
void clean_volume_creator_manager(GGEMSVolumeCreatorManager* volume_creator_manager);
GGEMSBox* create_box(float width, float height, float depth, char const* unit);
void delete_box(GGEMSBox* box);
void set_position_box(GGEMSBox* box, float pos_x, float pos_y, float pos_z, char const* unit);
void set_material_box(GGEMSBox* box, char const* material);
void set_label_value_box(GGEMSBox* box, float label_value);
void initialize_box(GGEMSBox* box);
void draw_box(GGEMSBox* box);
GGEMSTube* create_tube(float radius_x, float radius_y, float height, char const* unit);
void delete_tube(GGEMSTube* tube);
void set_position_tube(GGEMSTube* tube, float pos_x, float pos_y, float pos_z, char const* unit);
void set_material_tube(GGEMSTube* tube, char const* material);
void set_label_value_tube(GGEMSTube* tube, float label_value);
void initialize_tube(GGEMSTube* tube);
void draw_tube(GGEMSTube* tube);
GGEMSSphere* create_sphere(float radius, char const* unit);
void delete_sphere(GGEMSSphere* sphere);
void set_position_sphere(GGEMSSphere* sphere, float pos_x, float pos_y, float pos_z, char const* unit);
void set_material_sphere(GGEMSSphere* sphere, char const* material);
void set_label_value_sphere(GGEMSSphere* sphere, float label_value);
void initialize_sphere(GGEMSSphere* sphere);
void draw_sphere(GGEMSSphere* sphere);

/* CT acquisition in one launch (captured as a hipGraph): projection p turns
 * the X-ray source and every CT system by (rx[p], ry[p], rz[p]) [unit] on
 * top of their placement - the matrices SetRotation would give - and
 * transports n_particles[p] primaries; each CT system saves a stack of n
 * projections at its save() basename (DimSize cols rows n).  The images
 * equal those of n consecutive runs with those placements. */
void run_acquisition_ggems(GGEMS* ggems, size_t n, const float* rx, const float* ry, const float* rz,
                           char const* unit, const uint64_t* n_particles);
/* the same over several phantoms in one launch: projection p sees the CT
 * systems and only phantom navigator phantoms[p] (its rank among the
 * voxelized and meshed phantoms in creation order; -1: every phantom).
 * Equal to the consecutive runs of a world holding that one phantom. */
void run_acquisition_phantoms_ggems(GGEMS* ggems, size_t n, const float* rx, const float* ry, const float* rz,
                                    char const* unit, const uint64_t* n_particles, const int32_t* phantoms);
/* replays the last acquisition (its random streams continue) and saves the accumulated stacks */
void repeat_acquisition_ggems(GGEMS* ggems);
/* times trial runs of n_trial primaries over launch shapes and keeps the
 * fastest; call after initialize_ggems and before any run (it restores the
 * seeded streams and zeroes every tally) */
void autotune_ggems(GGEMS* ggems, uint64_t n_trial);
char const* get_last_error_ggems(void);
/* deletes every registered navigator and source and restores the process-table defaults */
void clean_safely_ggems(void);
/* wall time of the transport of the last run_ggems (ms) and its engine counters:
 * out[8] = primaries, voxel steps, voxel interactions, detector steps,
 * detector interactions, navigator hops, mesh steps, mesh interactions */
double get_run_time_ms_ggems(GGEMS* ggems);
/* the engine context (ggh_context*, include/ggems_engine.h) of an initialised GGEMS object */
void* get_engine_context_ggems(GGEMS* ggems);
void get_engine_stats_ggems(GGEMS* ggems, uint64_t* out8);
/* raw cross-section table row of a material: nbins values, mm^-1 */
void get_cross_section_table_ggems_cross_sections(GGEMSCrossSections* cross_sections, char const* process_name,
                                                  char const* material_name, float* out, size_t nbins);
/* energy grid of the cross-section tables (nbins values, MeV) */
void get_energy_bins_ggems_cross_sections(GGEMSCrossSections* cross_sections, float* out, size_t nbins);
/* photon energy cut of a material once initialised (MeV) */
float get_photon_cut_ggems_materials(GGEMSMaterials* materials, char const* material_name);

#ifdef __cplusplus
}
#endif
#endif
