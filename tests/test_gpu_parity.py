"""Parity of the HIP library (through the C ABI) with the CPU oracle.

Bar: bit-exact block tables, refinement maps and allocation indices; FP32 fields bit-exact
against the oracle's FP32 instantiation (same float operation sequence); FP32 fields within
1e-5 relative L2 of the FP64 oracle per step (same input state), with the drift after N
independent steps reported. At the full benchmark sizes, size-independent properties
(determinism, table invariants, finite fields, the projection reducing divergence).
"""
import sys

import numpy as np
import pytest

import dcgrid_amd
import oracle
from dcgrid_amd import scenes
from dcgrid_amd._abi import (STAGE_APPLY_PRESSURE, STAGE_DIVERGENCE, STAGE_JACOBI, STAGE_JACOBI_INV,
                             STAGE_PROLONGATE, STAGE_RESTRICT, STAGE_VORTICITY, hash_map)
from parity_util import TABLE_KEYS, compare_tables, rel_l2

pytestmark = pytest.mark.gpu

NF = np.uint64(0xFFFFFFFFFFFFFFFF)
FIELDS_DC = ("density", "velocity", "temperature", "pressure", "t_pressure", "divergence", "fluidity")
FIELDS_U = ("density", "velocity", "temperature", "pressure", "t_pressure", "divergence", "fluidity")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bitwise(a, b, mask=None):
    a, b = np.asarray(a), np.asarray(b)
    if mask is not None:
        a, b = a[mask], b[mask]
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def allocated_cells(sim):
    """Cells of allocated blocks: cells of free blocks are never read and hold no defined value
    (the reference copies its scratch buffer over them, fluid_simulation_dcgrid.cu:265)."""
    inf = sim.info()
    t = sim.export_tables()
    return np.repeat(t["block_levels"] != 0xFF, inf.block_width ** 3)


def dc_workload(name):
    if name == "ref64_bw4":
        p = scenes.reference_params(64)
        return "dcgrid", (64, 64, 64), p, 1500, False
    if name == "ref64_bw4_solids":
        p = scenes.reference_params(64)
        p.enable_additional_solids = 1
        return "dcgrid", (64, 64, 64), p, 2000, False
    if name == "ref64_bw8":
        p = scenes.reference_params(64)
        p.block_width = 8
        return "dcgrid", (64, 64, 64), p, 300, False
    if name == "cfg1_64":
        w = scenes.config(1, scale=4)
        return w.kind, w.size, w.params, w.max_num_blocks, True
    if name == "cfg1_128":
        w = scenes.config(1, scale=2)
        return w.kind, w.size, w.params, w.max_num_blocks, True
    if name == "cfg2_128":
        w = scenes.config(2, scale=4)
        return w.kind, w.size, w.params, w.max_num_blocks, True
    if name == "cfg4_128":  # terrain + moist atmosphere + wind, all levels dense
        w = scenes.config(4, scale=8)
        return w.kind, w.size, w.params, w.max_num_blocks, True
    if name == "cfg4_256":  # level 0 has 8192 dense blocks: the multi-block-per-iteration sweeps
        w = scenes.config(4, scale=4)
        return w.kind, w.size, w.params, w.max_num_blocks, True
    if name == "cfg4_512":  # 65536 dense blocks at level 0, 8192 at level 1: four per iteration
        w = scenes.config(4, scale=2)
        return w.kind, w.size, w.params, w.max_num_blocks, True
    if name == "cfg1_256_b12000":  # a sparse level of 7392 blocks and a dense one of 4096: two per iteration
        w = scenes.config(1)
        return w.kind, w.size, w.params, 12000, True
    if name == "cfg1_256_b20000":  # a budget of >= 16384 blocks: the two-cell stencils, no fluidity
        w = scenes.config(1)
        return w.kind, w.size, w.params, 20000, True
    if name == "cfg4_128_sparse":  # the same scene with sparse finer levels (moves, refinement)
        w = scenes.config(4, scale=8)
        w.params.score_mode = scenes.SCORE_GEOMETRIC
        w.params.refine_threshold = 1e-4
        return w.kind, w.size, w.params, 16 + 128 + 256, True
    raise KeyError(name)


def pair(name, precision=32):
    kind, size, p, b, seed = dc_workload(name)
    say(f"[case] {name} create")
    g = dcgrid_amd.Simulation(kind, size, p, b)
    c = oracle.OracleSim(kind, size, p, b, precision=precision)
    if seed:
        g.seed_scene()
        c.seed_scene()
    return g, c


def assert_tables_equal(g, c):
    tg, tc = g.export_tables(), c.export_tables()
    assert compare_tables(tg, tc) == []
    assert hash_map(tg, g.info()) == hash_map(tc, c.info())
    ig, ic = g.info(), c.info()
    for k in ("block_loads", "move_limit"):
        assert ig.levels_list(k) == ic.levels_list(k)


DC_CASES = ["ref64_bw4", "ref64_bw4_solids", "ref64_bw8", "cfg1_64", "cfg1_128", "cfg2_128", "cfg4_128",
            "cfg4_128_sparse"]


@pytest.mark.parametrize("name", ["cfg4_256", "cfg4_512", "cfg1_256_b12000", "cfg1_256_b20000"])
def test_dcgrid_large_level_steps_bitwise(name):
    """Levels with at least two (sixteen) rounds of the launch grid take the Jacobi variant with
    two (four) blocks per iteration: config 4 at 256x64x256 has 8192 dense blocks at level 0, at
    512x128x512 65536 at level 0 (four) and 8192 at level 1; the config 1 scene with a 12000-block
    budget has a sparse level 0 of 7392 blocks and a dense level 1 of 4096. Budgets of 16384
    blocks and more (config 4 at 512x128x512 with fluidity, config 1 with 20000 blocks without)
    run the divergence, gradient and feature kernels with two cells per thread. Two steps
    bit-exact against the FP32 oracle."""
    g, c = pair(name)
    for step in range(2):
        say(f"[case] {name} step {step}")
        g.step(1)
        c.step(1)
        assert_tables_equal(g, c)
        m = allocated_cells(c)
        for f in FIELDS_DC + (("vapor",) if g.params.moist else ()):
            assert bitwise(g.export_field(f), c.export_field(f), m), (name, step, f)
    g.close()
    c.close()


@pytest.mark.parametrize("name", DC_CASES)
def test_dcgrid_steps_bitwise_vs_fp32_oracle(name):
    g, c = pair(name)
    assert_tables_equal(g, c)  # after reset + init (5 adaptations)
    for step in range(4):
        say(f"[case] {name} step {step}")
        for stage in ("advect_velocity", "adapt_topology", "project", "advect_density"):
            getattr(g, stage)()
            getattr(c, stage)()
            if stage == "adapt_topology":
                assert_tables_equal(g, c)
        m = allocated_cells(c)
        for f in FIELDS_DC + (("vapor",) if g.params.moist else ()):
            assert bitwise(g.export_field(f), c.export_field(f), m), (name, step, f)
        assert g.debug_stats() == c.debug_stats()


@pytest.mark.parametrize("name", ["cfg1_64", "ref64_bw4_solids"])
def test_dcgrid_project_local_bitwise(name):
    """projectLocal (fluid_simulation_dcgrid.cu:296-311): divergence, project_local_pairs Jacobi
    pairs on every level from the top without prolongation, gradient subtraction and velocity
    restriction; after two reference steps, the GPU's velocity, pressure and divergence equal the
    FP32 oracle's bit for bit (a sparse scene, and BW 4 with sphere solids)."""
    g, c = pair(name)
    g.step(2)
    c.step(2)
    assert_tables_equal(g, c)
    for rep in range(2):
        say(f"[case] {name} project_local {rep}")
        g.project_local()
        c.project_local()
        m = allocated_cells(c)
        for f in ("velocity", "pressure", "t_pressure", "divergence"):
            assert bitwise(g.export_field(f), c.export_field(f), m), (name, rep, f)
    g.close()
    c.close()


@pytest.mark.parametrize("name", ["ref64_bw4", "ref64_bw4_solids", "cfg1_128"])
def test_dcgrid_fp64_per_step(name):
    """From the same input state, one GPU step is within 1e-5 relative L2 of the FP64 oracle."""
    g, c = pair(name, precision=64)
    worst = 0.0
    for step in range(3):
        say(f"[case] {name} fp64 step {step}")
        # same input: copy the GPU state into the FP64 oracle (topology is identical by parity)
        for f in ("density", "velocity", "temperature"):
            c.import_field(f, g.export_field(f))
        g.step(1)
        c.step(1)
        assert compare_tables(g.export_tables(), c.export_tables()) == []
        m = allocated_cells(c)
        for f in ("density", "velocity", "pressure"):
            r = rel_l2(g.export_field(f)[m], c.export_field_f64(f)[m])
            worst = max(worst, r)
            assert r <= 1e-5, (name, step, f, r)
    say(f"[result] {name}: worst per-step rel L2 vs FP64 = {worst:.3e}")


def test_dcgrid_fp64_drift():
    """Independent runs: FP32 GPU vs FP64 oracle after N steps (geometric topology, so both
    evolve the same tables). The drift is reported; it must stay far below the field scale."""
    g, c = pair("ref64_bw4", precision=64)
    out = []
    for n in range(1, 11):
        g.step(1)
        c.step(1)
        if n in (1, 5, 10):
            m = allocated_cells(c)
            r = rel_l2(g.export_field("velocity")[m], c.export_field_f64("velocity")[m])
            out.append((n, r))
    say("[result] drift FP32 GPU vs FP64 oracle (velocity rel L2):", out)
    assert out[0][1] <= 1e-5
    assert out[-1][1] <= 5e-6  # measured 3.6e-7 after 10 steps


@pytest.mark.parametrize("name", ["ref64_bw4", "cfg1_64", "ref64_bw8"])
def test_dcgrid_stage_parity_random_inputs(name):
    """Identical random fields into both; every stage bit-exact."""
    g, c = pair(name)
    inf = g.info()
    rng = np.random.default_rng(12)
    n = g.field_size("density")
    data = {
        "density": rng.random(n, np.float32),
        "temperature": rng.random(n, np.float32),
        "velocity": rng.standard_normal((n, 3)).astype(np.float32),
        "pressure": rng.standard_normal(n).astype(np.float32),
        "t_pressure": rng.standard_normal(n).astype(np.float32),
        "divergence": rng.standard_normal(n).astype(np.float32),
    }
    stages = [(STAGE_DIVERGENCE, 0), (STAGE_RESTRICT, "divergence"), (STAGE_RESTRICT, "velocity"),
              (STAGE_RESTRICT, "density"), (STAGE_VORTICITY, 0), (STAGE_APPLY_PRESSURE, 0)]
    for lvl in range(inf.levels):
        stages += [(STAGE_JACOBI, lvl), (STAGE_JACOBI_INV, lvl)]
    for lvl in range(inf.levels - 1):
        stages += [(STAGE_PROLONGATE, lvl)]
    for st, arg in stages:
        say(f"[case] {name} stage {st} {arg}")
        for s in (g, c):
            for f, v in data.items():
                s.import_field(f, v)
            s.run_stage(st, arg)
        m = allocated_cells(c)
        for f in ("density", "velocity", "pressure", "t_pressure", "divergence", "vorticity"):
            assert bitwise(g.export_field(f), c.export_field(f), m), (name, st, arg, f)


UNIFORM_CASES = ["cfg0_32", "ref32", "ref32_solids", "ref48x32x40"]


def uniform_workload(name):
    if name == "cfg0_32":
        w = scenes.config(0, scale=2)
        return w.size, w.params, True
    if name == "ref32":
        return (32, 32, 32), scenes.reference_params(32, kind="uniform"), False
    if name == "ref32_solids":
        p = scenes.reference_params(32, kind="uniform")
        p.enable_additional_solids = 1
        return (32, 32, 32), p, False
    if name == "ref48x32x40":
        return (48, 32, 40), scenes.reference_params(48, 32, 40, kind="uniform"), False
    raise KeyError(name)


@pytest.mark.parametrize("name", UNIFORM_CASES)
def test_uniform_steps_bitwise(name):
    size, p, seed = uniform_workload(name)
    say(f"[case] uniform {name}")
    g = dcgrid_amd.Simulation("uniform", size, p)
    c = oracle.OracleSim("uniform", size, p, precision=32)
    if seed:
        g.seed_scene()
        c.seed_scene()
    for f in FIELDS_U:
        assert bitwise(g.export_field(f), c.export_field(f)), (name, "init", f)
    for step in range(3):
        say(f"[case] uniform {name} step {step}")
        g.step(1)
        c.step(1)
        for f in FIELDS_U:
            assert bitwise(g.export_field(f), c.export_field(f)), (name, step, f)
    assert g.debug_stats() == c.debug_stats()
    # the other projection schedule too
    g.project_local()
    c.project_local()
    assert bitwise(g.export_field("velocity"), c.export_field("velocity"))


def test_uniform_fp64_per_step():
    size, p, _ = uniform_workload("cfg0_32")
    g = dcgrid_amd.Simulation("uniform", size, p)
    c = oracle.OracleSim("uniform", size, p, precision=64)
    g.seed_scene()
    worst = 0.0
    for step in range(3):
        for f in ("density", "velocity", "temperature"):
            c.import_field(f, g.export_field(f))
        g.step(1)
        c.step(1)
        for f in ("density", "velocity", "pressure"):
            r = rel_l2(g.export_field(f), c.export_field_f64(f))
            worst = max(worst, r)
            assert r <= 1e-5, (step, f, r)
    say(f"[result] uniform cfg0: worst per-step rel L2 vs FP64 = {worst:.3e}")


def test_gpu_deterministic():
    runs = []
    for _ in range(2):
        kind, size, p, b, _ = dc_workload("cfg1_128")
        g = dcgrid_amd.Simulation(kind, size, p, b)
        g.seed_scene()
        g.step(5)
        m = allocated_cells(g)
        runs.append((g.export_tables(), g.export_field("velocity")[m], g.export_field("density")[m]))
        g.close()
    assert compare_tables(runs[0][0], runs[1][0], TABLE_KEYS) == []
    assert bitwise(runs[0][1], runs[1][1]) and bitwise(runs[0][2], runs[1][2])


def check_table_invariants(t, inf):
    bw, L = inf.block_width, inf.levels
    lev, pos, par, chi = t["block_levels"], t["block_positions"], t["parent_indices"], t["child_indices"]
    loads = inf.levels_list("block_loads")
    offs = inf.levels_list("level_offsets")
    for level in range(L):
        idx = np.nonzero(lev == level)[0]
        assert len(idx) == loads[level]
        assert len(idx) == 0 or (idx[0] == offs[level] and idx[-1] == offs[level] + loads[level] - 1)
    alloc = np.nonzero((lev != 0xFF) & (lev < L - 1))[0]
    ps = par[alloc].astype(np.int64)
    assert np.all(chi[ps] == alloc)
    pb = ps // 8
    d = np.stack([(ps >> 2) & 1, (ps >> 1) & 1, ps & 1], 1)
    assert np.array_equal(pos[alloc], 2 * pos[pb] + bw * d)
    assert np.all(lev[pb] == lev[alloc] + 1)


@pytest.mark.parametrize("cfg", [1, 2])
def test_full_size_properties(cfg):
    """At the benchmark sizes: determinism, table invariants, finite fields, and the projection
    reducing the divergence."""
    w = scenes.config(cfg)
    say(f"[case] full size config {cfg} {w.size} budget {w.max_num_blocks}")
    results = []
    for rep in range(2):
        g = dcgrid_amd.Simulation(w.kind, w.size, w.params, w.max_num_blocks)
        g.seed_scene()
        g.step(8)
        g.advect_velocity()
        g.adapt_topology()
        g.run_stage(STAGE_DIVERGENCE)
        div_before = np.linalg.norm(g.export_field("divergence"))
        g.project()
        g.run_stage(STAGE_DIVERGENCE)
        div_after = np.linalg.norm(g.export_field("divergence"))
        t = g.export_tables()
        inf = g.info()
        check_table_invariants(t, inf)
        v = g.export_field("velocity")[allocated_cells(g)]
        assert np.all(np.isfinite(v))
        assert div_after < div_before
        results.append((t, v))
        say(f"[result] config {cfg}: loads {inf.levels_list('block_loads')} |div| {div_before:.4e} -> {div_after:.4e}")
        g.close()
    assert compare_tables(results[0][0], results[1][0]) == []
    assert bitwise(results[0][1], results[1][1])


def test_gpu_error_paths():
    p = scenes.reference_params(256)
    p.max_levels = 3
    with pytest.raises(dcgrid_amd.DCGError) as e:
        dcgrid_amd.Simulation("dcgrid", (256, 256, 256), p, 4096)  # BW 4, 3 levels: 4096 blocks are too few
    assert e.value.code == -2
    p.block_width = 6
    with pytest.raises(dcgrid_amd.DCGError) as e:
        dcgrid_amd.Simulation("dcgrid", (256, 256, 256), p, 4096)
    assert e.value.code == -4
    w = scenes.config(1, scale=4)
    g = dcgrid_amd.Simulation(w.kind, w.size, w.params, w.max_num_blocks)
    with pytest.raises(dcgrid_amd.DCGError):
        g.run_stage(STAGE_JACOBI, 7)  # no such level
    with pytest.raises(ValueError):
        g.import_field("density", np.zeros(3, np.float32))
