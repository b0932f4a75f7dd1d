/*
 * oracle.h — C API of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is a CPU restatement of the reference's simulation
 * step (wouterraateland/dcgrid, the src/dcgrid and src/uniformgrid kernels and src/utils/sim_utils.cu),
 * templated on the field precision (float or double). It exists to check the HIP library;
 * only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and never
 * as the thing measured or shipped.
 *
 * The API mirrors include/dcgrid_amd.h with an `orc_` prefix and an extra `precision`
 * argument (32 or 64) at creation, so tests drive both sides with identical calls.
 */
#ifndef DCGRID_ORACLE_H
#define DCGRID_ORACLE_H

#include "../include/dcgrid_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

int orc_create_uniform(const int32_t size[3], const dcg_params* p, int32_t precision,
                       orc_sim** out);
int orc_create_dcgrid(const int32_t size[3], uint64_t max_num_blocks, const dcg_params* p,
                      int32_t precision, orc_sim** out);
void orc_destroy(orc_sim* s);
int orc_set_params(orc_sim* s, const dcg_params* p);
int orc_init(orc_sim* s);
int orc_reset(orc_sim* s);
int orc_adapt_topology(orc_sim* s);
int orc_advect_velocity(orc_sim* s);
int orc_project(orc_sim* s);
int orc_project_local(orc_sim* s);
int orc_advect_density(orc_sim* s);
int orc_project_solve(orc_sim* s);
int orc_solve_report(orc_sim* s, dcg_solve_info* out);
int orc_step(orc_sim* s, int32_t n);
int orc_debug_stats(orc_sim* s, double* out);
/* Ray-marched image (src/raymarching.cuh), RGBA float per pixel, row-major. */
int orc_render(orc_sim* s, const dcg_render_params* r, float* out);
int orc_seed_scene(orc_sim* s);
int orc_info_get(orc_sim* s, dcg_info* out);
int orc_field_size(orc_sim* s, int32_t field, uint64_t* nfloats);
/* Fields are exported/imported as float (FP64 instances round on export) or as double. */
int orc_export_field(orc_sim* s, int32_t field, float* dst, uint64_t nfloats);
int orc_export_field_f64(orc_sim* s, int32_t field, double* dst, uint64_t nfloats);
int orc_import_field(orc_sim* s, int32_t field, const float* src, uint64_t nfloats);
int orc_import_field_f64(orc_sim* s, int32_t field, const double* src, uint64_t nfloats);
int orc_export_tables(orc_sim* s, dcg_tables* t);
/* Differentiable rollout (extension): record n steps from the current state (DCGrid: on a
 * frozen block topology, no adaptTopology), then map the adjoint of the final state (density,
 * velocity xyz-interleaved, temperature; gtem/otem may be NULL) to the gradient with respect to
 * the recorded initial state. FP64 in the FP64 instance: the reference-logic gradient. */
int orc_tape_forward(orc_sim* s, int32_t n);
int orc_tape_backward(orc_sim* s, const double* gden, const double* gvel, const double* gtem, double* oden,
                      double* ovel, double* otem);
/* One stage of the step (DCG_STAGE_* of dcgrid_amd.h). */
int orc_run_stage(orc_sim* s, int32_t stage, int32_t arg);
/* getBlockIndexDeep (dcgrid_utils.cuh:201-232) on the oracle's hash tables; -1 if none. */
int64_t orc_lookup(orc_sim* s, int32_t x, int32_t y, int32_t z, int32_t level);
/* Test hooks (DCGrid). orc_scores: the subblock scores (8B + 1 floats, the last one the
 * -FLT_MAX pad the block score's ninth term reads) and block scores (B floats) that the next
 * adaptTopology starts from (dcgrid_adaptation.cu:10-63). orc_import_tables: replace the block
 * tables (reference semantics, every member required), the per-level loads and move limits. */
int orc_scores(orc_sim* s, float* sub, float* blk);
int orc_import_tables(orc_sim* s, const dcg_tables* t, const uint64_t* loads, const uint64_t* move_limits);
/* Level/budget planner of the DCGrid constructor only (no allocation). */
int orc_plan(const int32_t size[3], uint64_t max_num_blocks, const dcg_params* p, dcg_info* out);
/* Number of OpenMP threads used by the oracle (0 = runtime default). */
void orc_set_threads(int32_t n);
int32_t orc_get_threads(void);
const char* orc_last_error(const orc_sim* s);

/* Scalar helpers exposed for closed-form tests of the restated index algebra. */
uint32_t orc_bit_spread3(uint32_t v);                             /* dcgrid_utils.cuh:83-90 */
uint32_t orc_grid_hash(int32_t x, int32_t y, int32_t z, int32_t bw); /* dcgrid_utils.cuh:92-96 */
uint32_t orc_local_index(int32_t x, int32_t y, int32_t z, int32_t bw);
uint64_t orc_mipmap_cells(uint64_t w, uint64_t h, uint64_t d);   /* grid_math.cuh:24-30 */
uint64_t orc_mipmap_idx(int32_t x, int32_t y, int32_t z, uint64_t w, uint64_t h, uint64_t d,
                        uint64_t scale);                          /* grid_math.cuh:43-52 */
float orc_value_noise(float x, float y, float z, uint32_t seed);
float orc_hash01(uint32_t seed, int32_t x, int32_t y, int32_t z, int32_t w);
/* Moist atmosphere and terrain (extension). */
float orc_exp_det(float x);
float orc_qsat(const dcg_params* p, float t);
float orc_terrain_height(const dcg_params* p, float x, float z);
float orc_cell_fluidity(const dcg_params* p, int32_t x, int32_t y, int32_t z, int32_t scale);
void orc_saturation_adjust(const dcg_params* p, float* qv, float* qc, float* t);
/* Field storage formats (extension): value after a store in format `storage` (DCG_STORAGE_*). */
void orc_round_storage(const float* in, float* out, uint64_t n, int32_t storage);

#ifdef __cplusplus
}
#endif

#endif
