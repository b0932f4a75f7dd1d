"""Seeded synthetic scenes.

`reference_params` restates SimParams::defaultParams (src/data/sim_params.cpp:4-17) with
rdx = 1/dx (src/main.cpp:17); `Scenes::Default` / `Scenes::SPGridCmp` are src/data/scenes.h:13-32.
The numbered workloads are the five benchmark configurations:

  0  uniform 64^3 buoyant smoke plume, 100 Jacobi sweeps, 50 steps
  1  adaptive DCGrid 256^3, 8^3 blocks, 3 levels, 4096-block budget, feature refinement, 200 steps
  2  adaptive DCGrid 512^3, 8^3 blocks, 4 levels, 32768-block budget, 100 steps
  3  uniform 128^3 (differentiable control; forward part)
  4  terrain cloud scene on a DCGrid 1024x256x1024 (y up), 8^3 blocks, 4 levels, all dense at
     its budget: fBm terrain, moist atmosphere with condensation, ambient wind
"""
from dataclasses import dataclass

from ._abi import (Params, PROJECT_LOCAL, PROJECT_MULTILEVEL, SCORE_FEATURES, SCORE_GEOMETRIC,
                   SOLID_TERRAIN, SOURCE_NONE, SOURCE_SPHERE_BUOYANT)

SEED_BASE = 20261202


def reference_params(gx=128, gy=None, gz=None, kind="dcgrid"):
    """SimParams::defaultParams with the grid size replaced; dx = 10000/gx as in scenes.h."""
    gy = gx if gy is None else gy
    gz = gx if gz is None else gz
    p = Params()
    p.gx, p.gy, p.gz = gx, gy, gz
    p.dx = 10000.0 / gx
    p.rdx = 1.0 / p.dx
    p.dt = 3.0
    p.velocity_emission_rate = 150.0
    p.density_emission_rate = 0.002
    p.emission_radius = 750.0
    p.enable_additional_solids = 0
    p.block_width = 4
    p.max_levels = 0
    p.project_mode = PROJECT_MULTILEVEL
    if kind == "dcgrid":
        # fluid_simulation_dcgrid.cu:274,284,301
        p.project_top_pairs, p.project_level_pairs, p.project_local_pairs = 5, 5, 10
    else:
        # fluid_simulation_uniform.cu:103,115,129
        p.project_top_pairs, p.project_level_pairs, p.project_local_pairs = 2, 1, 5
    p.score_mode = SCORE_GEOMETRIC
    p.refine_threshold = 1e-4  # fluid_simulation_dcgrid.cu:461
    p.grad_threshold = 0.05
    p.vort_threshold = 0.5
    p.source_mode = SOURCE_NONE
    p.t_ambient = 0.0
    p.noise_period = 32.0
    p.seed = SEED_BASE
    return p


@dataclass
class Workload:
    name: str
    kind: str
    size: tuple
    params: Params
    steps: int
    max_num_blocks: int = 0


def _plume(p, center, radius, noise_vel=0.0):
    p.dt = 0.1
    p.dx = 1.0
    p.rdx = 1.0
    p.velocity_emission_rate = 0.0
    p.density_emission_rate = 0.0
    p.emission_radius = 0.0  # closed box: no inlet
    p.source_mode = SOURCE_SPHERE_BUOYANT
    p.src_center[0], p.src_center[1], p.src_center[2] = center
    p.src_radius = radius
    p.src_density = 1.0
    p.src_temperature = 1.0
    p.src_noise = 0.01
    p.buoy_alpha = 0.05
    p.buoy_beta = 0.1
    p.t_ambient = 0.0
    p.init_velocity_noise = noise_vel
    return p


def terrain_scene(p, scale=1):
    """Terrain cloud scene (extension): a 10 m grid (times `scale`), 1 s steps, a 5-octave fBm
    heightfield of 60 cells (600 m) with a 256-cell lattice period, a moist layer of 1 km at
    90 % relative humidity under a 6.5 K/km lapse rate from 20 C at the ground, saturation
    adjustment with latent heating, expansion cooling of rising air (9.8 K/km), cloud-water
    loading and thermal buoyancy, and a 2 m/s ambient wind in a seeded horizontal direction."""
    import math
    p.dt = 1.0
    p.dx = 10.0 * scale
    p.rdx = 1.0 / p.dx
    p.velocity_emission_rate = 0.0
    p.density_emission_rate = 0.0
    p.emission_radius = 0.0
    p.source_mode = SOURCE_NONE
    p.enable_additional_solids = 1
    p.solid_mode = SOLID_TERRAIN
    p.terrain_height = 60.0 / scale
    p.terrain_period = 256.0 / scale
    p.moist = 1
    p.t_ambient = 20.0
    p.lapse_rate = 0.0065 * p.dx          # K per finest cell
    p.qsat_scale = 3.8e-3                  # kg/kg at 0 C (Magnus form)
    p.humidity = 0.9
    p.moist_layer = 100.0 / scale          # 1 km
    p.cond_rate = 0.5
    p.latent_heat = 2.5e6 / 1004.0         # L / c_p: K per kg/kg condensed
    p.adiabatic_rate = 0.0098              # K per m of ascent
    p.buoy_alpha = 9.81                    # cloud-water loading
    p.buoy_beta = 9.81 / 293.15            # thermal expansion
    theta = 2.0 * math.pi * ((SEED_BASE + 4) % 360) / 360.0
    p.wind[0], p.wind[1], p.wind[2] = 2.0 * math.cos(theta), 0.0, 2.0 * math.sin(theta)
    p.init_velocity_noise = 0.2
    p.noise_period = 64.0 / scale
    return p


def config(i, scale=1):
    """Workload of benchmark configuration i. `scale` shrinks the domain for quick tests."""
    if i == 0:
        n = 64 // scale
        p = reference_params(n, kind="uniform")
        _plume(p, (n / 2, n / 8, n / 2), 8.0 / scale)
        p.seed = SEED_BASE + 0
        p.project_mode = PROJECT_LOCAL
        p.project_local_pairs = 50  # "100 Jacobi iterations" = 100 sweeps = 50 pairs
        return Workload("uniform64_plume", "uniform", (n, n, n), p, 50)
    if i == 1:
        n = 256 // scale
        p = reference_params(n)
        _plume(p, (n / 2, n / 8, n / 2), 32.0 / scale, noise_vel=0.2)
        p.seed = SEED_BASE + 1
        p.block_width = 8
        p.max_levels = 3
        p.score_mode = SCORE_FEATURES
        p.refine_threshold = 1.0
        budget = 4096 // (scale ** 3) if scale > 1 else 4096
        return Workload("dcgrid256_bw8_l3_b4096", "dcgrid", (n, n, n), p, 200, budget)
    if i == 2:
        n = 512 // scale
        p = reference_params(n)
        _plume(p, (n / 2, n / 8, n / 2), 64.0 / scale, noise_vel=0.2)
        p.seed = SEED_BASE + 2
        p.block_width = 8
        p.max_levels = 4
        p.score_mode = SCORE_FEATURES
        p.refine_threshold = 1.0
        budget = 32768 // (scale ** 3) if scale > 1 else 32768
        return Workload("dcgrid512_bw8_l4_b32768", "dcgrid", (n, n, n), p, 100, budget)
    if i == 3:
        n = 128 // scale
        p = reference_params(n, kind="uniform")
        p.dt, p.dx, p.rdx = 0.1, 1.0, 1.0
        p.emission_radius = 0.0
        p.velocity_emission_rate = 0.0
        p.density_emission_rate = 0.0
        p.seed = SEED_BASE + 3
        return Workload("uniform128_control", "uniform", (n, n, n), p, 64)
    if i == 4:
        gx, gz, gy = 1024 // scale, 1024 // scale, 256 // scale
        p = reference_params(gx, gy, gz)
        terrain_scene(p, scale)
        p.seed = SEED_BASE + 4
        p.block_width = 8
        # four levels need the height to be a multiple of 8 * 2^3 cells
        p.max_levels = 4 if gy % 64 == 0 else 3
        p.score_mode = SCORE_FEATURES
        p.refine_threshold = 1.0
        L = p.max_levels
        full = sum((gx // (8 << l)) * (gy // (8 << l)) * (gz // (8 << l)) for l in range(L))
        return Workload("dcgrid1024x256x1024_bw8_l4_terrain", "dcgrid", (gx, gy, gz), p, 300, full)
    raise ValueError(i)


def control_inputs(n, seed=SEED_BASE + 3):
    """Seeded inputs of the differentiable density-control workload on an n^3 uniform grid
    (x-fastest cell order): initial velocity N(0, 0.1^2) per component, initial density a
    centred Gaussian blob (sigma 16 n/128 cells, peak 1), and the target density: 4 isotropic
    Gaussian blobs (sigma 6 n/128 cells, peak 1) at seeded centres. Returns float32
    (density0[N], velocity0[N,3], target[N])."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    N = n ** 3
    vel = rng.normal(0.0, 0.1, size=(N, 3)).astype(np.float32)
    c = np.arange(n) + 0.5
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")

    def blob(cx, cy, cz, sigma):
        return np.exp(-((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / (2.0 * sigma ** 2))

    den = blob(n / 2, n / 2, n / 2, 16.0 * n / 128).reshape(-1).astype(np.float32)
    centres = rng.uniform(0.25 * n, 0.75 * n, size=(4, 3))
    tgt = np.zeros((n, n, n))
    for cx, cy, cz in centres:
        tgt = np.maximum(tgt, blob(cx, cy, cz, 6.0 * n / 128))
    return den, vel, tgt.reshape(-1).astype(np.float32)
