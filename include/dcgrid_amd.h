/*
 * dcgrid_amd.h — C ABI of the MI355X (gfx950) DCGrid / uniform-grid simulation step.
 *
 * This is the drop-in boundary for the reference's abstract step interface
 * `FluidSimulation` (src/fluid_simulation.h:4-26) and its two back-ends
 * `FluidSimulationDCGrid` (src/dcgrid/fluid_simulation_dcgrid.h:5-61) and
 * `FluidSimulationUniform` (src/uniformgrid/fluid_simulation_uniform.h:5-39),
 * plus the global parameter upload `copySimParamsToDevice` (src/utils/sim_utils.h:6).
 *
 * Conventions
 *  - Plain C: opaque handle, POD parameters, caller-owned host buffers, int status codes.
 *  - No entry point ever calls exit(); failures return a negative DCG_E_* code and
 *    dcg_last_error() describes the last failure of that handle.
 *  - One host thread per handle. Calls are ordered on the handle's own HIP stream; calls
 *    that return host data synchronise that stream.
 *  - Parameters are per handle (the reference keeps them in one process-global
 *    __constant__ block, src/utils/sim_utils.cu:6-9).
 */
#ifndef DCGRID_AMD_H
#define DCGRID_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DCG_ABI_VERSION 2
#define DCG_MAX_LEVELS 16

/* Status codes. */
enum {
  DCG_OK = 0,
  DCG_E_ARG = -1,         /* bad argument (null pointer, unknown field, bad size) */
  DCG_E_BUDGET = -2,      /* block budget too small: src/dcgrid/fluid_simulation_dcgrid.cu:35-53 */
  DCG_E_HIP = -3,         /* a HIP runtime call failed */
  DCG_E_UNSUPPORTED = -4, /* configuration outside what this build implements */
  DCG_E_HASH_FULL = -5,   /* per-level hash table full: src/dcgrid/dcgrid_utils.cuh:138 */
  DCG_E_NOMEM = -6        /* device allocation failed */
};

/* Grid kinds reported by dcg_info. */
enum { DCG_KIND_UNIFORM = 0, DCG_KIND_DCGRID = 1 };

/* Field identifiers for export/import. Sizes are reported by dcg_field_size (in floats).
 * DCGrid fields are indexed by global cell index block*BW^3 + Morton(local) (reference layout,
 * src/dcgrid/dcgrid.h:29-43); uniform fields are x-fastest (src/utils/grid_math.cuh:10),
 * pressure/fluidity/divergence of the uniform grid carry the whole mip pyramid
 * (src/utils/grid_math.cuh:24-52). VELOCITY and VORTICITY are float3 AoS (3 floats per cell). */
enum {
  DCG_FIELD_DENSITY = 0,
  DCG_FIELD_VELOCITY = 1,
  DCG_FIELD_FLUIDITY = 2,
  DCG_FIELD_PRESSURE = 3,
  DCG_FIELD_T_PRESSURE = 4,
  DCG_FIELD_DIVERGENCE = 5,
  DCG_FIELD_TEMPERATURE = 6,
  DCG_FIELD_VORTICITY = 7,
  DCG_FIELD_VAPOR = 8, /* water vapor of the moist atmosphere (extension) */
  DCG_NUM_FIELDS = 9
};

/* Projection schedule used by dcg_step. DCG_PROJECT_SOLVE (extension) solves, level by level
 * from the coarsest, the fixed point of the Jacobi sweeps of DCG_PROJECT_MULTILEVEL to the
 * relative residual solve_tol (dcg_project_solve). */
enum { DCG_PROJECT_MULTILEVEL = 0, DCG_PROJECT_LOCAL = 1, DCG_PROJECT_SOLVE = 2 };

/* Arithmetic of DCG_PROJECT_SOLVE: FP64 residual / norms / correction update around FP32
 * multigrid cycles with red-black Gauss-Seidel smoothing (MIXED), or everything in FP32 or
 * in FP64. */
enum { DCG_SOLVE_MIXED = 0, DCG_SOLVE_FP32 = 1, DCG_SOLVE_FP64 = 2 };

/* Topology score used by dcg_adapt_topology. */
enum {
  DCG_SCORE_GEOMETRIC = 0, /* reference: src/dcgrid/dcgrid_adaptation.cu:26-34 */
  DCG_SCORE_FEATURES = 1   /* density-gradient / vorticity thresholds (extension) */
};

/* Sources. */
enum { DCG_SOURCE_NONE = 0, DCG_SOURCE_SPHERE_BUOYANT = 1 };

/* Solid geometry when enable_additional_solids is set: the reference's sphere
 * (src/sdf.cuh:8-18) or a seeded fBm heightfield (extension). */
enum { DCG_SOLID_SPHERE = 0, DCG_SOLID_TERRAIN = 1 };

/* Storage format of the DCGrid fields in HBM (extension): every field value is rounded to nearest
 * even when a kernel stores it and widened to FP32 when loaded; arithmetic stays FP32. Fields are
 * exported and imported as FP32 (exact widening, rounding on import). */
enum { DCG_STORAGE_FP32 = 0, DCG_STORAGE_FP16 = 1, DCG_STORAGE_BF16 = 2 };

/*
 * Simulation parameters. The first block equals the simulation fields of the reference's
 * SimParams (src/data/sim_params.h:14-30; its render fields are in dcg_render_params). Everything
 * after `enable_additional_solids` is a knob the reference hard-codes, or an extension.
 */
typedef struct dcg_params {
  float dt;
  int32_t gx, gy, gz;
  float dx, rdx; /* rdx must equal 1/dx (the reference sets it in src/main.cpp:17) */
  float velocity_emission_rate;
  float density_emission_rate;
  float emission_radius;
  int32_t enable_additional_solids;

  /* Block grid shape (DCGrid only). */
  int32_t block_width; /* 4 = reference layout (src/dcgrid/dcgrid.h:52), or 8 */
  int32_t max_levels;  /* 0 = reference rule (fluid_simulation_dcgrid.cu:12-22); else a cap */

  /* Projection schedule (counts of jacobi/jacobi_inv pairs). */
  int32_t project_mode;        /* DCG_PROJECT_* used by dcg_step */
  int32_t project_top_pairs;   /* DCGrid 5 (fluid_simulation_dcgrid.cu:274), uniform 2 (uniform.cu:103) */
  int32_t project_level_pairs; /* DCGrid 5 (:284), uniform 1 (uniform.cu:115) */
  int32_t project_local_pairs; /* DCGrid 10 (:301), uniform 5 (uniform.cu:129) */

  /* Topology scoring (DCGrid only). */
  int32_t score_mode;     /* DCG_SCORE_* */
  float refine_threshold; /* 1e-4 in the reference (fluid_simulation_dcgrid.cu:461) */
  float grad_threshold;   /* DCG_SCORE_FEATURES: |grad density| threshold */
  float vort_threshold;   /* DCG_SCORE_FEATURES: |vorticity| threshold */

  /* Sources (extension; none exist in the reference). */
  int32_t source_mode;   /* DCG_SOURCE_* */
  float src_center[3];   /* in finest-level cell units */
  float src_radius;      /* in finest-level cell units */
  float src_density;     /* density inside the sphere */
  float src_temperature; /* temperature inside the sphere (plus noise) */
  float src_noise;       /* amplitude of the seeded uniform temperature noise */
  float buoy_alpha;      /* buoyancy: u_y += dt*(beta*(T - t_ambient) - alpha*density) */
  float buoy_beta;
  float t_ambient;
  float init_velocity_noise; /* amplitude of the seeded 3-octave value-noise initial velocity */
  float noise_period;        /* lattice period of the first octave, finest cells */
  uint32_t seed;

  /* DCG_PROJECT_SOLVE (extension, DCGrid). */
  int32_t solve_precision;  /* DCG_SOLVE_* */
  float solve_tol;          /* per-level relative residual |r| / |r(0)| to reach (1e-8; <= 0: 1e-8) */
  int32_t solve_max_cycles; /* refinement cycles per level before giving up (50; <= 0: 50) */
  int32_t solve_smooth;     /* smoothing passes before and after each coarse correction, each
                               four red-black Gauss-Seidel sweeps per block colour
                               (1; <= 0: 1; at most 2) */

  /* Terrain and moist atmosphere (extension). */
  int32_t solid_mode;     /* DCG_SOLID_* */
  float terrain_height;   /* heightfield maximum, finest cells */
  float terrain_period;   /* lattice period of the first of 5 fBm octaves, finest cells */
  int32_t moist;          /* 1: vapor field, saturation adjustment with latent heating, ambient
                             temperature t_ambient - lapse_rate * height, buoyancy from it */
  float lapse_rate;       /* ambient temperature drop per finest cell of height */
  float qsat_scale;       /* saturation vapor q_s(T) = qsat_scale * exp(17.27 T / (T + 237.3)) */
  float humidity;         /* relative humidity of the ambient moist layer (initial and inflow) */
  float moist_layer;      /* height of the moist layer, finest cells */
  float cond_rate;        /* fraction of the super- (or under-) saturation converted per step */
  float latent_heat;      /* temperature change per unit of condensed water */
  float wind[3];          /* ambient wind, world units per second (initial field, lateral inflow) */
  float adiabatic_rate;   /* cooling of rising air per world unit of ascent (moist only) */
  int32_t storage;        /* DCG_STORAGE_* of the fields, fixed at creation (DCGrid only) */
  int32_t solve_warm;     /* DCG_PROJECT_SOLVE initial guess: 0 = zero on the top level and the
                             solved parent level prolonged below; 1 = the pressure of the last
                             projection on every block that kept its place (blocks moved or
                             created since are prolonged). The prolonged parent is the better
                             guess on dense scenes (config 4: 6 cycles per level against 16) */
  int32_t reserved[1];
} dcg_params;

/* Outcome of the last DCG_PROJECT_SOLVE projection, per level (0 finest). */
typedef struct dcg_solve_info {
  int32_t levels;
  int32_t cycles[DCG_MAX_LEVELS];          /* refinement cycles run on the level */
  double rel_residual[DCG_MAX_LEVELS];     /* |r| / |r(0)| reached */
  double initial_residual[DCG_MAX_LEVELS]; /* |r(0)|: residual of a zero level pressure */
  int32_t singular[DCG_MAX_LEVELS];        /* level covers the domain: pure Neumann, zero-mean */
  int32_t converged;                       /* every level reached solve_tol */
} dcg_solve_info;

/* Shape/state summary of a simulation. Per-level arrays are indexed by level (0 finest). */
typedef struct dcg_info {
  int32_t kind;        /* DCG_KIND_* */
  int32_t size[3];
  int32_t block_width; /* DCGrid */
  int32_t levels;      /* DCGrid levels or uniform mip levels */
  int32_t sparse_levels;
  int32_t apron_width; /* BW + 2 */
  uint64_t max_num_blocks;
  uint64_t num_cells;     /* cells of the field arrays (B*BW^3 or gx*gy*gz) */
  uint64_t active_cells;  /* sum over levels of block_loads*BW^3, or gx*gy*gz */
  uint64_t hash_table_size;
  uint64_t max_blocks_level[DCG_MAX_LEVELS];
  uint64_t full_blocks_level[DCG_MAX_LEVELS];
  uint64_t level_offsets[DCG_MAX_LEVELS];
  uint64_t block_loads[DCG_MAX_LEVELS];
  uint64_t move_limit[DCG_MAX_LEVELS];
  uint64_t hash_size_level[DCG_MAX_LEVELS];
  uint64_t hash_offset_level[DCG_MAX_LEVELS];
} dcg_info;

/* Block tables in the reference's semantics (src/dcgrid/dcgrid.h:5-27), for parity checks.
 * Caller-owned host buffers; a NULL member is skipped. notFound = UINT64_MAX.
 * cell_indices is the (BW+2)^3 apron table per block in [i][j][k] x-major order
 * (src/dcgrid/dcgrid.h:73). The hash is reported slot by slot; slot placement follows
 * insertion order and is only meaningful as a map (level, key) -> block. */
typedef struct dcg_tables {
  int32_t* block_positions; /* 3*B */
  uint8_t* block_levels;    /* B, 0xFF = free */
  uint8_t* block_flags;     /* B */
  uint64_t* parent_indices; /* B */
  uint64_t* child_indices;  /* 8*B */
  uint64_t* cell_indices;   /* B*(BW+2)^3 */
  uint32_t* hash_key;       /* hash_table_size */
  uint64_t* hash_val;       /* hash_table_size */
} dcg_tables;

typedef struct dcg_sim dcg_sim;

/* Reference defaults: src/data/sim_params.cpp:4-17, rdx = 1/dx (src/main.cpp:17). */
void dcg_default_params(dcg_params* p);

/* Replaces FluidSimulationUniform::FluidSimulationUniform (fluid_simulation_uniform.cu:6-56):
 * allocate, then reset() -> init(). */
int dcg_create_uniform(const int32_t size[3], const dcg_params* p, dcg_sim** out);

/* Replaces FluidSimulationDCGrid::FluidSimulationDCGrid(size, maxNumBlocks)
 * (fluid_simulation_dcgrid.cu:9-140): level/budget planning, allocation, reset() -> init(). */
int dcg_create_dcgrid(const int32_t size[3], uint64_t max_num_blocks, const dcg_params* p,
                      dcg_sim** out);

/* Replaces the destructors (fluid_simulation_dcgrid.cu:142-188, uniform.cu:58-73). */
void dcg_destroy(dcg_sim* s);

/* Replaces copySimParamsToDevice (src/utils/sim_utils.cu:7-9), per handle. Grid size,
 * block width and level cap are fixed at creation and must not change. */
int dcg_set_params(dcg_sim* s, const dcg_params* p);
int dcg_get_params(const dcg_sim* s, dcg_params* p);

/* FluidSimulation::init / reset (src/fluid_simulation.h:9-10). */
int dcg_init(dcg_sim* s);
int dcg_reset(dcg_sim* s);
/* FluidSimulation::adaptTopology (src/fluid_simulation.h:11). */
int dcg_adapt_topology(dcg_sim* s);
/* FluidSimulation::advectVelocity / project / projectLocal / advectDensity
 * (src/fluid_simulation.h:13-16). */
int dcg_advect_velocity(dcg_sim* s);
int dcg_project(dcg_sim* s);
int dcg_project_local(dcg_sim* s);
int dcg_advect_density(dcg_sim* s);
/* enable != 0: dcg_step replays pairs of steps through a HIP graph captured on first use (and
 * again after anything the recorded launches depend on changed: parameters, a single-stage call
 * that swapped field buffers). Per-kernel timing and DCG_PROJECT_SOLVE run eagerly. Off by
 * default. Extension; no counterpart in the reference. */
int dcg_set_graph(dcg_sim* s, int32_t enable);
/* Extension: the projection with the level systems solved to solve_tol (DCG_PROJECT_SOLVE),
 * and the report of the last such solve (synchronises). DCGrid only. */
int dcg_project_solve(dcg_sim* s);
int dcg_solve_report(dcg_sim* s, dcg_solve_info* out);
/* n steps in the order of src/simulation.cpp:104-111:
 * advectVelocity -> adaptTopology -> project (or projectLocal) -> advectDensity. */
int dcg_step(dcg_sim* s, int32_t n);
/* FluidSimulation::debugStats (src/fluid_simulation.h:22): DCGrid returns the residual sum
 * (dcgrid_structure.cu:224-251) of the last projection, uniform returns the smoke mass
 * (uniformgrid_structure.cu:33-42). Synchronises. */
int dcg_debug_stats(dcg_sim* s, double* out);

/* Sets the seeded initial condition selected by the parameters (sphere source, value-noise
 * velocity) on the current topology. Extension; no counterpart in the reference. */
int dcg_seed_scene(dcg_sim* s);

/* Active cells summed over the steps run by dcg_step since creation or the last reset (for a
 * DCGrid, the cells of all allocated blocks after each step's adaptTopology). Synchronises. */
int dcg_cell_steps(dcg_sim* s, uint64_t* total);

/* One stage of the step, for tests and custom schedules. `arg` is the level for the per-level
 * stages and the field id for DCG_STAGE_RESTRICT (restricts that field on all levels). */
enum {
  DCG_STAGE_DIVERGENCE = 0,     /* k_dcgrid_calc_divergence / k_uniform_calc_divergence (no restriction) */
  DCG_STAGE_RESTRICT = 1,       /* accumulate<T> (DCGrid) / k_uniform_restrict of the divergence (uniform) */
  DCG_STAGE_JACOBI = 2,         /* one half-sweep pressure -> t_pressure on level `arg` */
  DCG_STAGE_JACOBI_INV = 3,     /* one half-sweep t_pressure -> pressure on level `arg` */
  DCG_STAGE_PROLONGATE = 4,     /* level `arg` from level arg+1 */
  DCG_STAGE_APPLY_PRESSURE = 5, /* gradient subtraction (no restriction) */
  DCG_STAGE_VORTICITY = 6       /* DCGrid only */
};
int dcg_run_stage(dcg_sim* s, int32_t stage, int32_t arg);

/* Waits for all work queued on the handle. */
int dcg_synchronize(dcg_sim* s);

int dcg_info_get(dcg_sim* s, dcg_info* out);
int dcg_field_size(dcg_sim* s, int32_t field, uint64_t* nfloats);
/* Host <-> device copies of one field (synchronising). */
int dcg_export_field(dcg_sim* s, int32_t field, float* dst, uint64_t nfloats);
int dcg_import_field(dcg_sim* s, int32_t field, const float* src, uint64_t nfloats);
/* The same copies with caller-owned DEVICE memory (e.g. torch.Tensor.data_ptr()); the caller
 * finishes its own work on that memory first; both synchronise the handle's stream. Uniform
 * grid, and DCGrid with FP32 field storage. Extension; no counterpart in the reference. */
int dcg_export_field_device(dcg_sim* s, int32_t field, float* dst_device, uint64_t nfloats);
int dcg_import_field_device(dcg_sim* s, int32_t field, const float* src_device, uint64_t nfloats);

/* Differentiable rollout (extension; the reference has no derivatives).
 * dcg_tape_forward runs `steps` steps from the current state and keeps a checkpoint of the
 * state (density, temperature, velocity) every `checkpoint_every` steps. Uniform grid: the steps
 * of dcg_step. DCGrid: the steps on a frozen block topology (advectVelocity, the Jacobi
 * projection, advectDensity; no adaptTopology), so the block tables and fluidity stay as they
 * are. dcg_tape_backward takes the adjoint of the final state (d loss / d density, velocity as
 * xyz-interleaved float3, temperature; g_temperature may be NULL for zero) and writes the
 * gradient with respect to the state at the start of the rollout (d_temperature may be NULL),
 * recomputing each segment forward from its checkpoint. All six pointers are caller-owned
 * device memory of N (3N for velocity) floats, N the field size of dcg_field_size. On the DCGrid
 * the gradient's entries for cells of free blocks are 0, and its sums are formed with FP32
 * atomics (reproducible to rounding, not bitwise). Returns DCG_E_UNSUPPORTED with the projection
 * solve (DCG_PROJECT_SOLVE), with the moist atmosphere, for DCGrid 16-bit field storage, or
 * (uniform grid) when a foot moves more than 4 cells in one step. Afterwards the live state is
 * the rollout's final state. Synchronises. */
int dcg_tape_forward(dcg_sim* s, int32_t steps, int32_t checkpoint_every);
int dcg_tape_backward(dcg_sim* s, const float* g_density, const float* g_velocity, const float* g_temperature,
                      float* d_density, float* d_velocity, float* d_temperature);

/* Block tables in reference semantics (DCGrid only). */
int dcg_export_tables(dcg_sim* s, dcg_tables* t);

/* Rendering (FluidSimulation::render, src/fluid_simulation.h:19; k_dcgrid_render
 * src/dcgrid/dcgrid_rendering.cu:66-92, k_uniform_render src/uniformgrid/uniformgrid_rendering.cu,
 * ray marcher src/raymarching.cuh:63-275). The channels are RenderChannel
 * (src/data/sim_params.h:4-12); the render fields of SimParams (sim_params.h:32-44) and the
 * arguments of render() form dcg_render_params. The image is written as float4 RGBA per pixel,
 * pixel (x, y) at float offset 4 * (y * width + x) (the reference's surf2Dwrite at column x,
 * row y). dcg_render writes caller-owned device memory on the handle's stream;
 * dcg_render_host copies the image into caller-owned host memory and synchronises. */
enum {
  DCG_RENDER_FLUID = 0,      /* ray-marched smoke with shadows, floor and solids */
  DCG_RENDER_DENSITY = 1,    /* debug slice z = gz/2: density */
  DCG_RENDER_VELOCITY = 2,   /* debug slice: |velocity| */
  DCG_RENDER_FLUIDITY = 3,   /* debug slice: fluidity */
  DCG_RENDER_RESOLUTION = 4  /* debug slice: checkerboard of 4-cell squares */
};
typedef struct dcg_render_params {
  int32_t width, height;     /* resolution (render()'s int2) */
  float eye_origin[3];       /* world units */
  float eye_dir[3];          /* the view is rotated about x by (eye_dir.y, eye_dir.z) */
  float fov;
  float sun_dir[3];          /* unit vector towards the sun */
  int32_t channel;           /* DCG_RENDER_* */
  int32_t solids;            /* render_solids: floor and the solid scene */
  int32_t shadows;           /* render_shadows */
  int32_t precise;           /* render_precise: trilinear over the finest block's apron */
  float aa_samples;          /* samples per pixel side */
  float ambient;
  float background_color[3], floor_color[3], smoke_color[3], scene_color[3];
} dcg_render_params;
/* SimParams::defaultParams render fields (src/data/sim_params.cpp:19-32), a 256 x 256 view
 * from the middle of the domain's -z side, sun overhead-slanted. */
void dcg_default_render_params(const dcg_sim* s, dcg_render_params* r);
int dcg_render(dcg_sim* s, const dcg_render_params* r, float* rgba_device);
int dcg_render_host(dcg_sim* s, const dcg_render_params* r, float* rgba_host);
/* No counterpart in the reference: dcg_render_host for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, hosts that forbid hipMalloc). The
 * image is staged in a file-scope device array of at most 1024 x 1024 pixels (16 MiB per GPU,
 * held until the process ends) instead of per-call device memory; larger images are DCG_E_ARG.
 * It allocates nothing and gives the image dcg_render_host gives. Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (several handles or threads) race. */
int dcg_render_host_noalloc(dcg_sim* s, const dcg_render_params* r, float* rgba_host);

/* Versioned state of a simulation (extension: the reference keeps its state in device memory
 * only, src/dcgrid/dcgrid.h:5-43 and src/uniformgrid/uniformgrid.h:31-46). A state holds
 * everything later steps depend on, so a run resumed from it continues bit for bit as the run it
 * was saved from: the parameters, the DCGrid block tables (with the neighbour descriptors and the
 * hash verbatim) and device control words (block loads, move limits), and the fields in their
 * storage format. Counters (dcg_cell_steps), timing and the graph are not part of it.
 *
 * Layout (little endian, version DCG_STATE_VERSION):
 *   0   char[8]  "DCGSTATE"
 *   8   uint32   version
 *   12  uint32   header bytes = 72 + params bytes + 24 * sections
 *   16  int32    kind (DCG_KIND_*), storage (DCG_STORAGE_*), size[3], block_width, levels
 *               (the dcg_info values of the grid)
 *   44  uint32   flags (bit 0: the fluidity field was imported; bit 1: the pressure field holds
 *               a projection solve's solution, the next solve starts from it)
 *   48  uint32   params bytes (= sizeof(dcg_params)), 52 uint32 section count
 *   56  uint64   max_num_blocks, 64 uint64 num_cells (dcg_info)
 *   72  dcg_params
 *   then one 24-byte entry per section: uint32 id (DCG_STATE_*), uint32 element bytes,
 *   uint64 offset from the start, uint64 bytes; section data starts at 256-byte aligned offsets.
 * dcg_state_size gives the bytes of the grid's state; dcg_save_state writes it into a
 * caller-owned host buffer of at least that size; dcg_load_state restores a state saved from a
 * grid of the same kind, size, block width, levels, budget, storage and moist setting (else
 * DCG_E_ARG, before anything changes; an unknown version is DCG_E_UNSUPPORTED), parameters
 * included. Both synchronise. */
#define DCG_STATE_VERSION 1
enum {
  DCG_STATE_CONTROL = 1,       /* DCGrid: device control words, uint32 (loads, move limits, ...) */
  DCG_STATE_POSITIONS = 2,     /* DCGrid: block origin int32 x, y, z per block */
  DCG_STATE_LEVELS = 3,        /* DCGrid: uint8 per block, 0xFF free */
  DCG_STATE_FLAGS = 4,         /* DCGrid: uint32 per block */
  DCG_STATE_PARENTS = 5,       /* DCGrid: uint32 per block, 8*parent + subblock, 0xFFFFFFFF none */
  DCG_STATE_CHILDREN = 6,      /* DCGrid: uint32 per subblock (8 per block), 0xFFFFFFFF none */
  DCG_STATE_NEIGHBOURS = 7,    /* DCGrid: 27 int32x4 neighbour descriptors per block */
  DCG_STATE_HASH_KEYS = 8,     /* DCGrid: uint32 per hash slot */
  DCG_STATE_HASH_VALUES = 9,   /* DCGrid: uint32 per hash slot */
  DCG_STATE_FIELD = 16         /* + DCG_FIELD_*: the field in its storage format */
};
int dcg_state_size(dcg_sim* s, uint64_t* bytes);
int dcg_save_state(dcg_sim* s, void* dst, uint64_t bytes);
int dcg_load_state(dcg_sim* s, const void* src, uint64_t bytes);
/* A state buffer checked without a simulation (and without a device): header, section table
 * and, for a DCGrid, the block tables against the planner's sizes of the grid the header
 * describes: block loads within the level capacities (dense levels full), block levels
 * matching the loads, positions inside their level, parent, child, neighbour and hash links
 * pointing at allocated blocks of the right level. DCG_E_ARG on the first violation (message in
 * dcg_last_error(NULL)). dcg_load_state runs the same check before it copies anything. */
int dcg_state_check(const void* src, uint64_t bytes);

/* Per-kernel timing with HIP events recorded on the handle's stream around every launch of the
 * kernel classes selected by `mask` (bit k = kernel id k below; 0 disables). Setting the mask
 * clears the accumulated totals; dcg_timing_get synchronises and returns the total duration and
 * launch count of one class since then. Used by the benchmark for roofline figures. */
int dcg_set_timing(dcg_sim* s, int32_t mask);
int dcg_timing_get(dcg_sim* s, int32_t kernel, double* total_ms, uint64_t* launches);
/* Algorithmic HBM bytes of the launches timed in class `kernel` since dcg_set_timing, where the
 * library counts them itself (DCG_K_SOLVE_SMOOTH: the correction read and written and the
 * right-hand side read, per cell of the listed blocks; 0 for the other classes, whose work is a
 * function of the block loads the caller can read). */
int dcg_timing_bytes(dcg_sim* s, int32_t kernel, uint64_t* bytes);

/* Kernel ids for dcg_timing_get. */
enum {
  DCG_K_ADVECT_VELOCITY = 0,
  DCG_K_ADVECT_DENSITY = 1,
  DCG_K_DIVERGENCE = 2,
  DCG_K_JACOBI = 3,
  DCG_K_PROLONGATE = 4,
  DCG_K_APPLY_PRESSURE = 5,
  DCG_K_RESTRICT = 6,
  DCG_K_TOPOLOGY = 7,
  DCG_K_OTHER = 8,
  DCG_K_SOLVE_SMOOTH = 9, /* projection solve: block Gauss-Seidel smoothing of the block levels */
  DCG_K_SOLVE = 10,       /* projection solve: every other kernel of the solve */
  DCG_NUM_KERNEL_IDS = 11
};

const char* dcg_last_error(const dcg_sim* s);
int32_t dcg_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DCGRID_AMD_H */
