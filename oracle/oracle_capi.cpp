// oracle_capi.cpp — extern "C" surface of the CPU oracle (test infrastructure only; see oracle.h).
#include <omp.h>

#include "oracle_common.hpp"

namespace orc {
Sim* makeUniform(const int32_t size[3], const dcg_params& P, int precision, int* status, std::string* err);
Sim* makeDCGrid(const int32_t size[3], uint64_t maxBlocks, const dcg_params& P, int precision, int* status,
                std::string* err);
int planDCGrid(const int32_t size[3], uint64_t maxBlocks, const dcg_params& P, dcg_info* out, std::string* err);
}  // namespace orc

struct orc_sim {
  orc::Sim* impl;
};

static std::string g_create_error;

extern "C" {

int orc_create_uniform(const int32_t size[3], const dcg_params* p, int32_t precision, orc_sim** out) {
  if (!size || !p || !out || (precision != 32 && precision != 64)) return DCG_E_ARG;
  if (p->storage != DCG_STORAGE_FP32) {
    g_create_error = "16-bit field storage is implemented for the DCGrid";
    return DCG_E_UNSUPPORTED;
  }
  int st = 0;
  orc::Sim* s = orc::makeUniform(size, *p, precision, &st, &g_create_error);
  if (!s) return st;
  *out = new orc_sim{s};
  return DCG_OK;
}

int orc_create_dcgrid(const int32_t size[3], uint64_t max_num_blocks, const dcg_params* p, int32_t precision,
                      orc_sim** out) {
  if (!size || !p || !out || (precision != 32 && precision != 64)) return DCG_E_ARG;
  if (p->storage < DCG_STORAGE_FP32 || p->storage > DCG_STORAGE_BF16) return DCG_E_ARG;
  int st = 0;
  orc::Sim* s = orc::makeDCGrid(size, max_num_blocks, *p, precision, &st, &g_create_error);
  if (!s) return st;
  *out = new orc_sim{s};
  return DCG_OK;
}

void orc_destroy(orc_sim* s) {
  if (!s) return;
  delete s->impl;
  delete s;
}

#define IMPL(s) \
  if (!(s)) return DCG_E_ARG; \
  orc::Sim* I = (s)->impl

int orc_set_params(orc_sim* s, const dcg_params* p) {
  IMPL(s);
  if (!p) return DCG_E_ARG;
  if (p->gx != I->P.gx || p->gy != I->P.gy || p->gz != I->P.gz || p->block_width != I->P.block_width ||
      p->max_levels != I->P.max_levels || p->moist != I->P.moist || p->storage != I->P.storage)
    return I->fail(DCG_E_ARG,
                   "grid size, block width, level cap, moist atmosphere and field storage are fixed at creation");
  I->P = *p;
  return DCG_OK;
}
int orc_init(orc_sim* s) {
  IMPL(s);
  return I->init();
}
int orc_reset(orc_sim* s) {
  IMPL(s);
  return I->reset();
}
int orc_adapt_topology(orc_sim* s) {
  IMPL(s);
  return I->adaptTopology();
}
int orc_advect_velocity(orc_sim* s) {
  IMPL(s);
  return I->advectVelocity();
}
int orc_project(orc_sim* s) {
  IMPL(s);
  return I->project();
}
int orc_project_local(orc_sim* s) {
  IMPL(s);
  return I->projectLocal();
}
int orc_project_solve(orc_sim* s) {
  IMPL(s);
  return I->projectSolve();
}
int orc_solve_report(orc_sim* s, dcg_solve_info* out) {
  IMPL(s);
  if (!out) return DCG_E_ARG;
  return I->solveReport(out);
}
int orc_advect_density(orc_sim* s) {
  IMPL(s);
  return I->advectDensity();
}
// Step order of src/simulation.cpp:104-111.
int orc_step(orc_sim* s, int32_t n) {
  IMPL(s);
  for (int i = 0; i < n; i++) {
    int e;
    if ((e = I->advectVelocity())) return e;
    if ((e = I->adaptTopology())) return e;
    if ((e = I->projectByMode())) return e;
    if ((e = I->advectDensity())) return e;
  }
  return DCG_OK;
}
int orc_render(orc_sim* s, const dcg_render_params* r, float* out) {
  IMPL(s);
  if (!r || !out || r->width <= 0 || r->height <= 0 || !(r->aa_samples >= 1.f && r->aa_samples <= 8.f))
    return DCG_E_ARG;
  return I->render(*r, out);
}
int orc_debug_stats(orc_sim* s, double* out) {
  IMPL(s);
  if (!out) return DCG_E_ARG;
  return I->debugStats(out);
}
int orc_seed_scene(orc_sim* s) {
  IMPL(s);
  return I->seedScene();
}
int orc_info_get(orc_sim* s, dcg_info* out) {
  IMPL(s);
  if (!out) return DCG_E_ARG;
  return I->info(out);
}
int orc_field_size(orc_sim* s, int32_t field, uint64_t* n) {
  IMPL(s);
  if (!n) return DCG_E_ARG;
  return I->fieldSize(field, n);
}
int orc_export_field_f64(orc_sim* s, int32_t field, double* dst, uint64_t n) {
  IMPL(s);
  if (!dst) return DCG_E_ARG;
  return I->exportField(field, dst, n);
}
int orc_export_field(orc_sim* s, int32_t field, float* dst, uint64_t n) {
  IMPL(s);
  if (!dst) return DCG_E_ARG;
  std::vector<double> tmp(n);
  int e = I->exportField(field, tmp.data(), n);
  if (e) return e;
  for (uint64_t i = 0; i < n; i++) dst[i] = (float)tmp[i];
  return DCG_OK;
}
int orc_import_field(orc_sim* s, int32_t field, const float* src, uint64_t n) {
  IMPL(s);
  if (!src) return DCG_E_ARG;
  std::vector<double> tmp(src, src + n);
  return I->importField(field, tmp.data(), n);
}
int orc_import_field_f64(orc_sim* s, int32_t field, const double* src, uint64_t n) {
  IMPL(s);
  if (!src) return DCG_E_ARG;
  return I->importField(field, src, n);
}
int orc_tape_forward(orc_sim* s, int32_t n) {
  IMPL(s);
  return I->tapeForward(n);
}
int orc_tape_backward(orc_sim* s, const double* gden, const double* gvel, const double* gtem, double* oden,
                      double* ovel, double* otem) {
  IMPL(s);
  if (!gden || !gvel || !oden || !ovel) return DCG_E_ARG;
  return I->tapeBackward(gden, gvel, gtem, oden, ovel, otem);
}
int orc_export_tables(orc_sim* s, dcg_tables* t) {
  IMPL(s);
  if (!t) return DCG_E_ARG;
  return I->exportTables(t);
}
int orc_run_stage(orc_sim* s, int32_t stage, int32_t arg) {
  IMPL(s);
  return I->runStage(stage, arg);
}
int64_t orc_lookup(orc_sim* s, int32_t x, int32_t y, int32_t z, int32_t level) {
  if (!s) return -1;
  return s->impl->lookup(x, y, z, level);
}
int orc_scores(orc_sim* s, float* sub, float* blk) {
  IMPL(s);
  if (!sub || !blk) return DCG_E_ARG;
  return I->scores(sub, blk);
}
int orc_import_tables(orc_sim* s, const dcg_tables* t, const uint64_t* loads, const uint64_t* move_limits) {
  IMPL(s);
  if (!t) return DCG_E_ARG;
  return I->importTables(t, loads, move_limits);
}
int orc_plan(const int32_t size[3], uint64_t max_num_blocks, const dcg_params* p, dcg_info* out) {
  if (!size || !p || !out) return DCG_E_ARG;
  return orc::planDCGrid(size, max_num_blocks, *p, out, &g_create_error);
}
void orc_set_threads(int32_t n) {
  if (n > 0) omp_set_num_threads(n);
}
int32_t orc_get_threads(void) { return omp_get_max_threads(); }
const char* orc_last_error(const orc_sim* s) { return s ? s->impl->err.c_str() : g_create_error.c_str(); }

uint32_t orc_bit_spread3(uint32_t v) { return orc::bitSpread3(v); }
uint32_t orc_grid_hash(int32_t x, int32_t y, int32_t z, int32_t bw) { return orc::gridHash(x, y, z, bw); }
uint32_t orc_local_index(int32_t x, int32_t y, int32_t z, int32_t bw) {
  return orc::localIndex(x, y, z, bw == 4 ? 2 : 3);
}
uint64_t orc_mipmap_cells(uint64_t w, uint64_t h, uint64_t d) { return orc::mipmapCells(w, h, d); }
uint64_t orc_mipmap_idx(int32_t x, int32_t y, int32_t z, uint64_t w, uint64_t h, uint64_t d, uint64_t scale) {
  return orc::mipmapIdx(x, y, z, w, h, d, scale);
}
float orc_value_noise(float x, float y, float z, uint32_t seed) { return orc::valueNoise(x, y, z, seed); }
float orc_hash01(uint32_t seed, int32_t x, int32_t y, int32_t z, int32_t w) { return orc::hash01(seed, x, y, z, w); }
float orc_exp_det(float x) { return orc::expDet(x); }
float orc_qsat(const dcg_params* p, float t) { return orc::qsat(*p, t); }
float orc_terrain_height(const dcg_params* p, float x, float z) { return orc::terrainHeight(*p, x, z); }
float orc_cell_fluidity(const dcg_params* p, int32_t x, int32_t y, int32_t z, int32_t scale) {
  return orc::cellFluidity(*p, x, y, z, scale);
}
void orc_saturation_adjust(const dcg_params* p, float* qv, float* qc, float* t) {
  orc::saturationAdjust<float>(*p, *qv, *qc, *t);
}
void orc_round_storage(const float* in, float* out, uint64_t n, int32_t storage) {
  for (uint64_t i = 0; i < n; i++) out[i] = orc::roundStorage(in[i], storage);
}

}  // extern "C"
