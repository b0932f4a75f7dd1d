// Drives the C++ drop-in classes the way src/simulation.cpp:17-35 and :93-115 drive the
// reference's: copySimParamsToDevice, construct, init, then advectVelocity / adaptTopology /
// project / advectDensity per frame, and debugStats at the end. Prints one "cells <n>" line and
// the value debugStats returned, or "error <code> <message>" and exits 3 when the library throws.
//
// usage: drop_in_main dcgrid|uniform <gx> <gy> <gz> <max_blocks> <steps> <block_width> <levels>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "dcgrid_amd_fluid_simulation.hpp"

struct int3 {
  int x, y, z;
};
struct int2 {
  int x, y;
};
struct float3 {
  float x, y, z;
};

// The reference's render(dest, resolution, eyeOrigin, eyeDir, fov, sunDir) call shape
// (src/simulation.cpp:130-136) compiles against the drop-in (dest: device RGBA buffer).
void renderFrame(dcgrid_amd::FluidSimulation& sim, float* rgbaDevice) {
  sim.render(rgbaDevice, int2{64, 48}, float3{5000.f, 6000.f, -9000.f}, float3{0.f, -.2f, .98f}, .7f,
             float3{.3f, .9f, -.3f});
}

int main(int argc, char** argv) {
  if (argc != 9) {
    std::fprintf(stderr, "usage: %s dcgrid|uniform gx gy gz max_blocks steps block_width levels\n", argv[0]);
    return 2;
  }
  const bool adaptive = std::strcmp(argv[1], "dcgrid") == 0;
  const int3 size{std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4])};
  const size_t maxNumBlocks = std::strtoull(argv[5], nullptr, 10);
  const int steps = std::atoi(argv[6]);

  dcg_params p;
  dcg_default_params(&p);
  p.gx = size.x;
  p.gy = size.y;
  p.gz = size.z;
  p.dx = 10000.f / size.x;
  p.rdx = 1.f / p.dx;
  p.block_width = std::atoi(argv[7]);
  p.max_levels = std::atoi(argv[8]);
  try {
    dcgrid_amd::copySimParamsToDevice(p);
    std::unique_ptr<dcgrid_amd::FluidSimulation> sim;
    if (adaptive)
      sim = std::make_unique<dcgrid_amd::FluidSimulationDCGrid>(size, maxNumBlocks);
    else
      sim = std::make_unique<dcgrid_amd::FluidSimulationUniform>(size);
    sim->init();
    for (int i = 0; i < steps; i++) {
      dcgrid_amd::copySimParamsToDevice(p);
      sim->advectVelocity();
      sim->adaptTopology();
      sim->project();
      sim->advectDensity();
    }
    dcg_info inf;
    dcg_info_get(sim->handle(), &inf);
    std::printf("cells %llu\n", (unsigned long long)inf.num_cells);
    sim->debugStats();
  } catch (const dcgrid_amd::Error& e) {
    std::printf("error %d %s\n", e.code, e.what());
    return 3;
  }
  return 0;
}
