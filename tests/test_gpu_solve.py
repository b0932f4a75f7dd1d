"""GPU projection solve (DCG_PROJECT_SOLVE) through the C ABI: every level's system (the fixed
point of the reference's Jacobi sweeps) solved to a 1e-8 relative residual by FP64 iterative
refinement around FP32 multigrid cycles, checked against the FP64 oracle's SOR solution of the
same system (solved to 1e-12), against the all-FP32 and all-FP64 variants, and against the
reference Jacobi sweep itself (a solution is its fixed point)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dcgrid_amd  # noqa: E402
from dcgrid_amd import scenes  # noqa: E402
from dcgrid_amd._abi import PROJECT_SOLVE, SOLVE_FP32, SOLVE_FP64, SOLVE_MIXED, STAGE_JACOBI  # noqa: E402
from parity_util import rel_l2  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def workload(name):
    if name == "cfg2_128":
        w = scenes.config(2, scale=4)
        return w.kind, w.size, w.params, w.max_num_blocks
    if name == "cfg1_128":
        w = scenes.config(1, scale=2)
        return w.kind, w.size, w.params, w.max_num_blocks
    if name == "ref64_bw4":
        p = scenes.reference_params(64)
        return "dcgrid", (64, 64, 64), p, 1500
    if name == "cfg4_128":  # terrain solids: fluidity in the divergence and the gradient
        w = scenes.config(4, scale=8)
        return w.kind, w.size, w.params, w.max_num_blocks
    raise KeyError(name)


def prepared(name, precision, steps=4):
    """GPU and FP64-oracle simulations after `steps` reference steps, the oracle's velocity
    replaced by the GPU's (identical tables, identical projection input)."""
    import oracle
    kind, size, p, b = workload(name)
    p = p.copy()
    p.solve_precision = precision
    p.solve_tol = 1e-8
    p.solve_max_cycles = 60
    say(f"[case] {name} precision={precision}")
    g = dcgrid_amd.Simulation(kind, size, p, b)
    c = oracle.OracleSim(kind, size, p, b, precision=64)
    if name != "ref64_bw4":
        g.seed_scene()
        c.seed_scene()
    g.step(steps)
    c.step(steps)
    assert all(np.array_equal(a, b_) for a, b_ in zip(
        (g.export_tables()[k] for k in ("block_levels", "parent_indices")),
        (c.export_tables()[k] for k in ("block_levels", "parent_indices"))))
    c.import_field("velocity", g.export_field("velocity"))
    return g, c


def level_cells(sim, level):
    inf = sim.info()
    t = sim.export_tables()
    return np.repeat(t["block_levels"] == level, inf.block_width ** 3)


@pytest.mark.parametrize("name", ["cfg2_128", "ref64_bw4", "cfg1_128", "cfg4_128"])
def test_mixed_solve_reaches_tolerance_and_matches_oracle(name):
    g, c = prepared(name, SOLVE_MIXED)
    g.project_solve()
    c.project_solve()
    rg, rc = g.solve_report(), c.solve_report()
    say("gpu", rg)
    say("oracle", {k: rc[k] for k in ("cycles", "rel_residual")})
    assert rg["converged"]
    assert max(rg["rel_residual"]) <= 1e-8
    assert rg["singular"] == rc["singular"]
    mask = (g.export_tables()["block_levels"] != 0xFF).repeat(g.info().block_width ** 3)
    pg, pc = g.export_field("pressure")[mask], c.export_field_f64("pressure")[mask]
    vg, vc = g.export_field("velocity")[mask], c.export_field_f64("velocity")[mask]
    say("pressure rel l2", rel_l2(pg, pc), "velocity rel l2", rel_l2(vg, vc))
    # measured: pressure 2.8e-8 / 8.7e-8 / 2.9e-8, velocity 1.0e-7 / 3.1e-8 / 3.5e-7
    assert rel_l2(pg, pc) <= 2e-7
    assert rel_l2(vg, vc) <= 1e-6


def test_fixed_point_of_the_reference_jacobi_sweep():
    """A converged pressure is (to FP32 rounding) unchanged by one Jacobi half-sweep of the
    reference schedule on every level (dcgrid_multigrid_solver.cu:5-41)."""
    g, _ = prepared("cfg2_128", SOLVE_MIXED)
    g.project_solve()
    # the projection changed the velocity, but the pressure and divergence are those solved
    inf = g.info()
    rep = g.solve_report()
    p = g.export_field("pressure")
    for level in range(inf.levels):
        g.run_stage(STAGE_JACOBI, level)
        tp = g.export_field("t_pressure")
        m = level_cells(g, level)
        d = tp[m].astype(np.float64) - p[m]
        scale = np.linalg.norm(p[m]) / np.sqrt(m.sum())
        if rep["singular"][level]:
            # the singular level was solved with its right-hand side projected to zero mean: the
            # sweep then shifts every cell by the same constant (the removed mean / 6)
            d = d - d.mean()
        err = np.sqrt(np.mean(d * d)) / scale
        say("level", level, "singular", rep["singular"][level], "jacobi change", err)
        assert err <= 3e-7  # measured about 6e-8 on every level


def test_full_size_config2_small_regions():
    """Configuration 2 at its full 512^3 size, where sparse levels' regions (the top-level blocks
    holding their cells) cover at most 1/16 of the top level, so their V-cycles end at the top
    level without the pyramid: every level still reaches 1e-8, and the solution is (to FP32
    rounding) a fixed point of the reference Jacobi half-sweep on every level."""
    w = scenes.config(2)
    p = w.params.copy()
    p.project_mode = PROJECT_SOLVE
    p.solve_precision = SOLVE_MIXED
    p.solve_tol = 1e-8
    say("[case] config 2 at 512^3, 6 solved steps")
    g = dcgrid_amd.Simulation(w.kind, w.size, p, w.max_num_blocks)
    g.seed_scene()
    g.step(6)
    inf = g.info()
    t = g.export_tables()
    lev, par = t["block_levels"], t["parent_indices"]
    L, top = inf.levels, inf.levels - 1
    full = [(512 // (8 << l)) ** 3 for l in range(L)]
    loads = list(inf.block_loads)[:L]
    none = np.uint64(0xFFFFFFFFFFFFFFFF)
    small = []
    for l in range(top):
        if loads[l] >= full[l]:
            continue  # dense level
        b = np.nonzero(lev == l)[0]
        for _ in range(top - l):
            b = (par[b] // 8).astype(np.int64)
        tops = np.unique(b)
        say("sparse level", l, "region covers", tops.size, "of", full[top], "top-level blocks")
        if tops.size * 16 <= full[top]:
            small.append(l)
    assert small, "no sparse level with a small region: the case does not exercise that cycle"
    assert not np.any(par[lev == top] != none)
    g.project_solve()
    rep = g.solve_report()
    say("cycles", rep["cycles"], "rel", rep["rel_residual"])
    assert rep["converged"] and all(r <= 1e-8 for r in rep["rel_residual"])
    pr = g.export_field("pressure")
    for level in range(L):
        g.run_stage(STAGE_JACOBI, level)
        tp = g.export_field("t_pressure")
        m = level_cells(g, level)
        d = tp[m].astype(np.float64) - pr[m]
        scale = np.linalg.norm(pr[m]) / np.sqrt(m.sum())
        if rep["singular"][level]:
            d = d - d.mean()
        err = np.sqrt(np.mean(d * d)) / scale
        say("level", level, "small region" if level in small else "", "jacobi change", err)
        assert err <= 3e-7
    g.close()


def test_precision_variants():
    """FP64 throughout also reaches 1e-8 and agrees with the mixed solve; FP32 throughout stops
    where its arithmetic runs out (reported, not converged at 1e-8) with a larger error."""
    res = {}
    for prec in (SOLVE_MIXED, SOLVE_FP64, SOLVE_FP32):
        g, c = prepared("cfg2_128", prec)
        g.project_solve()
        if prec == SOLVE_MIXED:
            c.project_solve()
            res["oracle"] = c.export_field_f64("pressure")
        res[prec] = (g.solve_report(), g.export_field("pressure"), g)
    mask = (res[SOLVE_MIXED][2].export_tables()["block_levels"] != 0xFF).repeat(512)
    ref = res["oracle"][mask]
    e_mixed = rel_l2(res[SOLVE_MIXED][1][mask], ref)
    e_64 = rel_l2(res[SOLVE_FP64][1][mask], ref)
    e_32 = rel_l2(res[SOLVE_FP32][1][mask], ref)
    say("errors vs oracle: mixed", e_mixed, "fp64", e_64, "fp32", e_32)
    say("fp32 report", res[SOLVE_FP32][0])
    assert res[SOLVE_FP64][0]["converged"]
    assert max(res[SOLVE_FP64][0]["rel_residual"]) <= 1e-8
    assert e_64 <= 2e-7 and e_mixed <= 2e-7  # measured 2.8e-8 (both)
    assert max(res[SOLVE_FP32][0]["rel_residual"]) > 1e-8 or e_32 <= 1e-5


def test_step_with_solve_mode():
    """dcg_step with project_mode = DCG_PROJECT_SOLVE: every step's solve converges."""
    kind, size, p, b = workload("cfg2_128")
    p = p.copy()
    p.project_mode = PROJECT_SOLVE
    g = dcgrid_amd.Simulation(kind, size, p, b)
    g.seed_scene()
    g.step(3)
    rep = g.solve_report()
    say("step report", rep)
    assert rep["converged"]
    assert np.all(np.isfinite(g.export_field("velocity")))
