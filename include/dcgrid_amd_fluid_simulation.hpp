// dcgrid_amd_fluid_simulation.hpp — C++ drop-in for the reference's step interface, implemented
// over the C ABI of include/dcgrid_amd.h (header-only; link libdcgrid_amd.so).
//
// Mirrors src/fluid_simulation.h:4-26 (abstract FluidSimulation with init/reset/adaptTopology/
// advectVelocity/project/projectLocal/advectDensity/debugStats and protected size/numCells),
// src/dcgrid/fluid_simulation_dcgrid.h:7 (FluidSimulationDCGrid(size, maxNumBlocks)),
// src/uniformgrid/fluid_simulation_uniform.h:7 (FluidSimulationUniform(size)) and
// src/utils/sim_utils.h:6 (copySimParamsToDevice) and render(), which writes float4 RGBA into a
// device buffer instead of a cudaSurfaceObject_t.
//
// Differences a caller sees:
//  * errors throw dcgrid_amd::Error instead of printing and calling exit()
//    (helper_cuda.h getLastCudaError, fluid_simulation_dcgrid.cu:35-53);
//  * copySimParamsToDevice stores the parameters process-wide like the reference's __constant__
//    block; every simulation applies the latest ones before its next call.
#pragma once

#include <cstdio>
#include <stdexcept>
#include <string>

#include "dcgrid_amd.h"

namespace dcgrid_amd {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

namespace detail {
inline dcg_params& globalParams() {
  static dcg_params p = [] {
    dcg_params d;
    dcg_default_params(&d);
    return d;
  }();
  return p;
}
inline unsigned& globalGeneration() {
  static unsigned g = 1;
  return g;
}
inline void check(int st, const dcg_sim* s, const char* what) {
  if (st != DCG_OK) throw Error(st, std::string(what) + ": " + dcg_last_error(s));
}
}  // namespace detail

// copySimParamsToDevice (src/utils/sim_utils.cu:7-9): process-wide parameters.
inline void copySimParamsToDevice(const dcg_params& p) {
  detail::globalParams() = p;
  detail::globalGeneration()++;
}

class FluidSimulation {
 public:
  virtual ~FluidSimulation() { dcg_destroy(sim_); }
  FluidSimulation(const FluidSimulation&) = delete;
  FluidSimulation& operator=(const FluidSimulation&) = delete;

  void init() { call(dcg_init, "init"); }
  void reset() { call(dcg_reset, "reset"); }
  void adaptTopology() { call(dcg_adapt_topology, "adaptTopology"); }
  void advectVelocity() { call(dcg_advect_velocity, "advectVelocity"); }
  void project() { call(dcg_project, "project"); }
  void projectLocal() { call(dcg_project_local, "projectLocal"); }
  void advectDensity() { call(dcg_advect_density, "advectDensity"); }
  // The reference prints "Residual: %g" (DCGrid) or "Smoke: %g" (uniform); the value is also returned.
  double debugStats() {
    sync();
    double v = 0;
    detail::check(dcg_debug_stats(sim_, &v), sim_, "debugStats");
    std::printf(kind_ == DCG_KIND_DCGRID ? "Residual: %g\n" : "Smoke: %g\n", v);
    return v;
  }
  // render(dest, resolution, eyeOrigin, eyeDir, fov, sunDir) (src/fluid_simulation.h:19): dest is
  // device memory of 4 * resolution.x * resolution.y floats; I2/F3 are any types with int x, y /
  // float x, y, z members (int2, float3). Render settings other than the arguments (channel,
  // shadows, solids, colours) come from dcg_default_render_params unless `settings` is given.
  template <class I2, class F3>
  void render(float* dest, const I2& resolution, const F3& eyeOrigin, const F3& eyeDir, float fov, const F3& sunDir,
              const dcg_render_params* settings = nullptr) {
    dcg_render_params r;
    if (settings)
      r = *settings;
    else
      dcg_default_render_params(sim_, &r);
    r.width = resolution.x;
    r.height = resolution.y;
    const float o[3] = {eyeOrigin.x, eyeOrigin.y, eyeOrigin.z}, d[3] = {eyeDir.x, eyeDir.y, eyeDir.z},
                sd[3] = {sunDir.x, sunDir.y, sunDir.z};
    for (int a = 0; a < 3; a++) {
      r.eye_origin[a] = o[a];
      r.eye_dir[a] = d[a];
      r.sun_dir[a] = sd[a];
    }
    r.fov = fov;
    detail::check(dcg_render(sim_, &r, dest), sim_, "render");
  }
  // Handle for everything else in the C ABI (export/import, tables, timing).
  dcg_sim* handle() const { return sim_; }

 protected:
  struct Int3 {
    int x, y, z;
  };
  template <class V>
  static dcg_params paramsFor(const V& size) {
    dcg_params p = detail::globalParams();
    p.gx = size.x;
    p.gy = size.y;
    p.gz = size.z;
    return p;
  }
  FluidSimulation() = default;
  void adopt(dcg_sim* s, int kind, Int3 sz) {
    sim_ = s;
    kind_ = kind;
    size = sz;
    dcg_info inf;
    detail::check(dcg_info_get(sim_, &inf), sim_, "info");
    numCells = inf.num_cells;
    generation_ = detail::globalGeneration();
  }
  Int3 size{0, 0, 0};
  size_t numCells = 0;

 private:
  void sync() {
    if (generation_ == detail::globalGeneration()) return;
    dcg_params p = detail::globalParams();
    dcg_params cur;
    dcg_get_params(sim_, &cur);
    // grid shape, moist atmosphere and field storage are fixed at construction; the rest follows
    // the upload
    p.gx = cur.gx;
    p.gy = cur.gy;
    p.gz = cur.gz;
    p.block_width = cur.block_width;
    p.max_levels = cur.max_levels;
    p.moist = cur.moist;
    p.storage = cur.storage;
    detail::check(dcg_set_params(sim_, &p), sim_, "copySimParamsToDevice");
    generation_ = detail::globalGeneration();
  }
  void call(int (*f)(dcg_sim*), const char* what) {
    sync();
    detail::check(f(sim_), sim_, what);
  }
  dcg_sim* sim_ = nullptr;
  int kind_ = DCG_KIND_UNIFORM;
  unsigned generation_ = 0;
};

// FluidSimulationDCGrid(const int3& size, const size_t& maxNumBlocks)
// (src/dcgrid/fluid_simulation_dcgrid.h:7); V is any type with int members x, y, z (e.g. int3).
class FluidSimulationDCGrid : public FluidSimulation {
 public:
  template <class V>
  FluidSimulationDCGrid(const V& size, const size_t& maxNumBlocks) {
    const dcg_params p = paramsFor(size);
    const int32_t sz[3] = {size.x, size.y, size.z};
    dcg_sim* s = nullptr;
    detail::check(dcg_create_dcgrid(sz, maxNumBlocks, &p, &s), nullptr, "FluidSimulationDCGrid");
    adopt(s, DCG_KIND_DCGRID, Int3{size.x, size.y, size.z});
  }
};

// FluidSimulationUniform(const int3& size) (src/uniformgrid/fluid_simulation_uniform.h:7).
class FluidSimulationUniform : public FluidSimulation {
 public:
  template <class V>
  explicit FluidSimulationUniform(const V& size) {
    const dcg_params p = paramsFor(size);
    const int32_t sz[3] = {size.x, size.y, size.z};
    dcg_sim* s = nullptr;
    detail::check(dcg_create_uniform(sz, &p, &s), nullptr, "FluidSimulationUniform");
    adopt(s, DCG_KIND_UNIFORM, Int3{size.x, size.y, size.z});
  }
};

}  // namespace dcgrid_amd
