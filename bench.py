#!/usr/bin/env python3
"""Benchmark of the DCGrid simulation step on one MI355X.

Workload (default, configuration 2): the mixed-precision adaptive smoke plume on a 512^3 virtual
domain of 8^3-cell blocks, 4 levels, a 32768-block budget, feature-driven refinement, the buoyant
sphere source and the seeded value-noise initial velocity. Fields are stored in BF16 (FP32
arithmetic); the projection of every step solves each level's pressure system to a 1e-8 relative
residual by FP64 iterative refinement around FP32 multigrid cycles. Block tables are rebuilt
every step. Synthetic, seeded inputs, resident in HBM before the timed region.

One step = advectVelocity -> adaptTopology -> project -> advectDensity (src/simulation.cpp:104-111)
through the C ABI. Prints one JSON line:
  value          active cell-steps/s (allocated cells summed over the timed steps / time)
  roofline       the step's dominant kernel with a byte model: algorithmic bytes over its
                 HIP-event-timed average launch duration, against the 8 TB/s HBM peak
  kernels        per kernel class: time per step, launches, algorithmic GB/s where defined
  solve          the same projection solved from a cold start in mixed precision, all-FP32 and
                 all-FP64: time, cycles per level, reached residual, pressure error against FP64
  cpu_baseline   the CPU oracle (FP32 restatement, OpenMP over all host cores) on a bounded
                 sample of the same scene

Other workloads: --config 0..4 (3 = the adjoint rollout), --project reference (the reference's
Jacobi schedule), --storage fp32/fp16/bf16, --jacobi-micro (uniform 512^3 Jacobi sweeps, a
working set that does not fit the 256 MB Infinity Cache), --adjoint (DCGrid configurations: the
differentiable rollout on a frozen block topology, backward against forward time).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "active cell-steps/s + %HBM peak; mixed-prec vs FP32 projection at 1e-8 resid"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


STORAGE_BYTES = {0: 4, 1: 2, 2: 2}  # dcg_params.storage -> bytes per stored scalar
STORAGE_NAMES = {"fp32": 0, "fp16": 1, "bf16": 2}
DTYPE = {0: "f32", 1: "f16", 2: "bf16"}


def jacobi_bytes(info, params):
    """Algorithmic bytes and launch count of the Jacobi half-sweeps of one reference project()
    call: each half-sweep reads p and div and writes p (three stored scalars per cell)."""
    S = STORAGE_BYTES[params.storage]
    bv = int(info.block_width) ** 3
    L = int(info.levels)
    total_bytes = 0
    launches = 0
    for level in range(L):
        pairs = params.project_top_pairs if level == L - 1 else params.project_level_pairs
        cells = int(info.block_loads[level]) * bv
        total_bytes += 2 * pairs * cells * 3 * S
        launches += 2 * pairs
    return total_bytes, launches


def per_cell_bytes(params, moist=False):
    """Algorithmic bytes per active cell of the step's field kernels (S = stored scalar, V = 3 S,
    F = fluidity when solids are on, k = advected scalars)."""
    S = STORAGE_BYTES[params.storage]
    V = 3 * S
    F = S if params.enable_additional_solids else 0
    src = params.source_mode != 0 or params.moist
    k = (3 if params.moist else 2) if src else 1
    return {
        "advect_velocity": 2 * V + F + (2 * S if src else 0),
        "advect_density": V + 2 * k * S + F,
        "divergence": V + F + S,
        "apply_pressure": 2 * V + F + S,
    }


def step_bytes(info, params, uniform=False):
    """Algorithmic HBM bytes of one reference-schedule step (per-stage compulsory field traffic)."""
    S = STORAGE_BYTES[params.storage]
    V = 3 * S
    pc = per_cell_bytes(params)
    if uniform:
        act = int(info.active_cells)
        sweeps = 2 * (params.project_local_pairs if params.project_mode == 1 else params.project_level_pairs)
        return act * (pc["advect_velocity"] + pc["advect_density"] + pc["divergence"] + pc["apply_pressure"] +
                      12 * sweeps)
    bv = int(info.block_width) ** 3
    L = int(info.levels)
    src = params.source_mode != 0
    k = 2 if src else 1
    act = sum(int(info.block_loads[l]) for l in range(L)) * bv
    below = sum(int(info.block_loads[l]) for l in range(L - 1)) * bv
    jb, _ = jacobi_bytes(info, params)
    b = act * (pc["advect_velocity"] + pc["divergence"] + pc["apply_pressure"] + pc["advect_density"])
    b += below * 9 * V // 8 * 2          # restrict u twice
    b += below * 9 * S // 8              # restrict div
    b += jb                              # Jacobi sweeps
    b += below * S * 9 // 8              # prolongate
    b += below * 9 * k * S // 8          # restrict scalars
    return b


def copy_probe(nbytes=1 << 30):
    """A plain device copy of known size (read nbytes, write nbytes), three times: the reference
    that rocprofv3's FETCH_SIZE / WRITE_SIZE counters are scaled by."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"[bench] copy probe: 3 x {nbytes} B copied")
    del x, y


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the rocprofv3 counter summary committed for this
    workload (profiles/traffic_<kernel>_<workload>.txt, 'key: value' lines), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.txt")
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    try:
        return int(float(kv["traffic_bytes_per_launch"]))
    except (KeyError, ValueError):
        return None


def make_sim(cls, w, params=None, **kw):
    p = w.params if params is None else params
    s = cls(w.kind, w.size, p, w.max_num_blocks if w.kind == "dcgrid" else None, **kw)
    s.seed_scene()
    return s


def cpu_baseline(w, params, warm_steps, budget_s):
    """The oracle (FP32) on the same seeded scene: warm_steps untimed, then steps until about
    budget_s seconds of CPU work; reports active cell-steps/s."""
    import oracle
    oracle.library()
    cores = oracle.get_threads()
    c = make_sim(oracle.OracleSim, w, params, precision=32)
    c.step(warm_steps)
    n = 0
    cells = 0
    t0 = time.perf_counter()
    while True:
        inf = c.info()
        act = int(inf.active_cells)
        c.step(1)
        n += 1
        cells += act
        if time.perf_counter() - t0 > budget_s or n >= 2000:
            break
    dt = time.perf_counter() - t0
    c.close()
    return {
        "value": cells / dt,
        "unit": "active cell-steps/s",
        "cores": cores,
        "kind": "port",
        "seconds": round(dt, 3),
        "cells_per_step": round(cells / max(n, 1)),
        "steps": n,
    }


def adjoint_bench(w, args):
    """Differentiable control workload (uniform 128^3, 64 steps, checkpoints every 8): time of
    the plain forward rollout, of the recorded forward (with checkpoints) and of the backward
    pass (segment recomputation + adjoint kernels); value = backward / plain forward, and the
    adjoint steps alone (backward minus the recorded forward) / plain forward beside it."""
    import torch
    import dcgrid_amd
    from dcgrid_amd import scenes
    n = w.size[0]
    steps, k = w.steps, 8
    den, vel, tgt = scenes.control_inputs(n)
    sim = dcgrid_amd.Simulation(w.kind, w.size, w.params)
    dev = torch.device("cuda")
    d0 = torch.as_tensor(den, device=dev)
    v0 = torch.as_tensor(vel, device=dev)
    target = torch.as_tensor(tgt, device=dev)
    N = d0.numel()

    def load():
        torch.cuda.synchronize()
        sim.import_field_device("density", d0.data_ptr(), N)
        sim.import_field_device("velocity", v0.data_ptr(), 3 * N)

    load()
    sim.step(2)  # warm-up
    sim.synchronize()
    load()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    t_fwd = time.perf_counter() - t0
    load()
    t0 = time.perf_counter()
    sim.tape_forward(steps, k)
    sim.synchronize()
    t_tape = time.perf_counter() - t0
    dN = torch.empty_like(d0)
    sim.export_field_device("density", dN.data_ptr(), N)
    g = (dN - target).contiguous()
    gv = torch.zeros_like(v0)
    od, ov = torch.empty_like(d0), torch.empty_like(v0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.tape_backward(g.data_ptr(), gv.data_ptr(), None, od.data_ptr(), ov.data_ptr(), None)
    t_bwd = time.perf_counter() - t0
    # every state kept (checkpoint interval 1: 20 B per cell and step, 2.7 GB here): the backward
    # pass re-runs no step
    load()
    sim.tape_forward(steps, 1)
    sim.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.tape_backward(g.data_ptr(), gv.data_ptr(), None, od.data_ptr(), ov.data_ptr(), None)
    t_bwd1 = time.perf_counter() - t0
    sim.close()
    return {
        "metric": "adjoint/forward time ratio",
        "value": round(t_bwd / t_fwd, 3),
        "unit": "ratio",
        "n_gpus": 1,
        "steps": steps,
        "higher_is_better": False,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian velocity, Gaussian blob density, 4-blob target)",
        "config": {"workload": w.name, "domain": list(w.size), "checkpoint_every": k},
        "forward_ms": round(1e3 * t_fwd, 3),
        "recorded_forward_ms": round(1e3 * t_tape, 3),
        "backward_ms": round(1e3 * t_bwd, 3),
        # the backward pass re-runs every checkpoint segment forward (about the recorded
        # forward's cost); what remains is the adjoint steps themselves
        "adjoint_steps_ms": round(1e3 * (t_bwd - t_tape), 3),
        "adjoint_step_ratio": round((t_bwd - t_tape) / t_fwd, 3),
        "all_states_kept": {"checkpoint_every": 1, "backward_ms": round(1e3 * t_bwd1, 3),
                            "ratio": round(t_bwd1 / t_fwd, 3)},
        "forward_cell_steps_per_s": round(N * steps / t_fwd, 1),
        "grad_norm_velocity": float(torch.linalg.vector_norm(ov)),
    }


def adjoint_dcgrid_bench(w, p, args):
    """Differentiable rollout on the DCGrid (frozen block topology): the workload's scene after
    its warm-up steps with adaptation, then `--adjoint-steps` steps with checkpoints every 4.
    Times the rollout recorded with only its end states (the plain forward), the recorded
    rollout with checkpoints, and the backward pass (segment recomputation + adjoint kernels)
    for the seed d(0.5 |rho_N|^2); value = backward / plain forward."""
    import torch
    import dcgrid_amd
    steps, k = args.adjoint_steps, 4
    sim = make_sim(dcgrid_amd.Simulation, w, p)
    log(f"[bench] adjoint: {args.warmup} adaptive warm-up steps")
    sim.step(args.warmup)
    sim.synchronize()
    N = sim.export_field("density").size
    dev = torch.device("cuda")
    d0, v0, t0s = (torch.empty(N, device=dev), torch.empty(N, 3, device=dev), torch.empty(N, device=dev))
    sim.export_field_device("density", d0.data_ptr(), N)
    sim.export_field_device("velocity", v0.data_ptr(), 3 * N)
    sim.export_field_device("temperature", t0s.data_ptr(), N)

    def load():
        torch.cuda.synchronize()
        sim.import_field_device("density", d0.data_ptr(), N)
        sim.import_field_device("velocity", v0.data_ptr(), 3 * N)
        sim.import_field_device("temperature", t0s.data_ptr(), N)

    times = {}
    for name, kk in (("warm", steps), ("plain", steps), ("tape", k)):
        load()
        t = time.perf_counter()
        sim.tape_forward(steps, kk)
        times[name] = time.perf_counter() - t
    dN = torch.empty_like(d0)
    sim.export_field_device("density", dN.data_ptr(), N)
    gv, od, ov, ot = torch.zeros_like(v0), torch.empty_like(d0), torch.empty_like(v0), torch.empty_like(d0)
    torch.cuda.synchronize()
    log(f"[bench] adjoint: backward over {steps} steps")
    t = time.perf_counter()
    sim.tape_backward(dN.data_ptr(), gv.data_ptr(), None, od.data_ptr(), ov.data_ptr(), ot.data_ptr())
    t_bwd = time.perf_counter() - t
    # every state kept (checkpoint interval 1): the backward pass re-runs no step
    load()
    sim.tape_forward(steps, 1)
    torch.cuda.synchronize()
    t = time.perf_counter()
    sim.tape_backward(dN.data_ptr(), gv.data_ptr(), None, od.data_ptr(), ov.data_ptr(), ot.data_ptr())
    t_bwd1 = time.perf_counter() - t
    cells = int(sim.info().active_cells)
    sim.close()
    return {
        "metric": "adjoint/forward time ratio",
        "value": round(t_bwd / times["plain"], 3),
        "unit": "ratio",
        "n_gpus": 1,
        "steps": steps,
        "higher_is_better": False,
        "dtype": "f32",
        "data": "synthetic (the seeded scene after its warm-up steps)",
        "config": {"workload": w.name, "domain": list(w.size), "budget": w.max_num_blocks,
                   "checkpoint_every": k, "frozen_topology": True, "active_cells": cells},
        "forward_ms": round(1e3 * times["plain"], 3),
        "recorded_forward_ms": round(1e3 * times["tape"], 3),
        "backward_ms": round(1e3 * t_bwd, 3),
        "adjoint_steps_ms": round(1e3 * (t_bwd - times["tape"]), 3),
        "adjoint_step_ratio": round((t_bwd - times["tape"]) / times["plain"], 3),
        "all_states_kept": {"checkpoint_every": 1, "backward_ms": round(1e3 * t_bwd1, 3),
                            "ratio": round(t_bwd1 / times["plain"], 3)},
        "grad_norm_velocity": float(torch.linalg.vector_norm(ov)),
    }


def solve_compare(sim, params, reps=3):
    """The projection of the simulation's current velocity, solved from a cold start (zero top
    level, prolongation below) to 1e-8 in mixed precision, all-FP32 and all-FP64: wall time
    (median of `reps`, each from the same velocity), cycles and residual per level, and the
    relative L2 distance of the pressure to the FP64 solution on the allocated cells."""
    import numpy as np
    from dcgrid_amd._abi import SOLVE_FP32, SOLVE_FP64, SOLVE_MIXED
    vel = sim.export_field("velocity")
    mask = (sim.export_tables()["block_levels"] != 0xFF).repeat(int(sim.info().block_width) ** 3)
    out = {}
    press = {}
    for name, prec in (("fp64", SOLVE_FP64), ("mixed", SOLVE_MIXED), ("fp32", SOLVE_FP32)):
        q = params.copy()
        q.solve_precision = prec
        q.solve_tol = 1e-8
        sim.set_params(q)
        times = []
        for _ in range(reps):
            sim.import_field("velocity", vel)
            sim.synchronize()
            t0 = time.perf_counter()
            sim.project_solve()
            sim.synchronize()
            times.append(1e3 * (time.perf_counter() - t0))
        rep = sim.solve_report()
        press[name] = sim.export_field("pressure")[mask].astype(np.float64)
        times.sort()
        out[name] = {"ms": round(times[len(times) // 2], 3), "cycles": rep["cycles"],
                     "rel_residual": [float(f"{r:.3e}") for r in rep["rel_residual"]],
                     "converged": bool(rep["converged"])}
        log(f"[bench] solve {name}: {out[name]}")
    ref = press["fp64"]
    for name in ("mixed", "fp32"):
        out[name]["pressure_rel_l2_vs_fp64"] = float(f"{np.linalg.norm(press[name] - ref) / np.linalg.norm(ref):.3e}")
    out["mixed_over_fp64_time"] = round(out["mixed"]["ms"] / out["fp64"]["ms"], 3)
    sim.set_params(params)
    sim.import_field("velocity", vel)
    return out


def jacobi_micro(args):
    """Jacobi half-sweeps of the uniform grid's finest level at 512^3 (134 M cells, 1.6 GB per
    sweep: beyond the 256 MB Infinity Cache), timed with HIP events; 12 B per cell per sweep."""
    import numpy as np
    import dcgrid_amd
    from dcgrid_amd import scenes
    from dcgrid_amd._abi import STAGE_JACOBI, STAGE_JACOBI_INV
    n = 512
    p = scenes.reference_params(n, kind="uniform")
    sim = dcgrid_amd.Simulation("uniform", (n, n, n), p)
    rng = np.random.default_rng(7)
    sim.import_field("divergence", rng.standard_normal(sim.field_size("divergence")).astype(np.float32))
    for k in range(4):
        sim.run_stage(STAGE_JACOBI if k % 2 == 0 else STAGE_JACOBI_INV, 0)
    sim.synchronize()
    sim.set_timing(["jacobi"])
    launches = 2 * max(args.steps // 2, 10)
    for k in range(launches):
        sim.run_stage(STAGE_JACOBI if k % 2 == 0 else STAGE_JACOBI_INV, 0)
    sim.synchronize()
    ms, cnt = sim.timing()["jacobi"]
    per_launch = 12 * n ** 3
    achieved = per_launch / (ms / cnt * 1e-3) / 1e9
    sim.close()
    return {"metric": "uniform Jacobi half-sweep bandwidth (cache-busting)", "value": round(achieved, 1),
            "unit": "GB/s", "n_gpus": 1, "steps": int(cnt), "higher_is_better": True, "dtype": "f32",
            "data": "synthetic (seeded normal divergence)",
            "config": {"workload": "uniform512_jacobi", "domain": [n, n, n]},
            "roofline": {"kernel": "k_u_jacobi", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": measured_traffic("uniform512_jacobi", "k_u_jacobi"),
                         "avg_launch_us": round(1e3 * ms / cnt, 3), "algorithmic_bytes_per_launch": per_launch}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # default: configuration 2's 100-step run (10 untimed + 90 timed)
    ap.add_argument("--steps", type=int, default=90)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--storage", choices=sorted(STORAGE_NAMES), default=None,
                    help="field storage (default: bf16 for configuration 2, else fp32)")
    ap.add_argument("--project", choices=["solve", "reference"], default=None,
                    help="solve: every level solved to 1e-8 (mixed precision, default for configurations 2 and "
                         "4); reference: the reference's Jacobi schedule (default for 0, 1)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=5,
                    help="steps of the separate pass that times every launch with HIP events")
    ap.add_argument("--no-solve-compare", action="store_true")
    ap.add_argument("--graph", action="store_true",
                    help="replay the timed steps in pairs from one captured HIP graph (reference schedule)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 1 GiB device-to-device copy first (calibrates rocprofv3 byte counters)")
    ap.add_argument("--jacobi-micro", action="store_true", help="uniform 512^3 Jacobi bandwidth only")
    ap.add_argument("--adjoint", action="store_true",
                    help="DCGrid configurations: time the differentiable rollout (reference projection)")
    ap.add_argument("--adjoint-steps", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus above 1 is not supported "
                                   "(the step has no multi-GPU path)"}))
        sys.exit(2)

    import dcgrid_amd
    from dcgrid_amd import scenes
    from dcgrid_amd._abi import PROJECT_SOLVE, SOLVE_MIXED
    if args.copy_probe:
        copy_probe()
    if args.jacobi_micro:
        print(json.dumps(jacobi_micro(args)), flush=True)
        return
    w = scenes.config(args.config)
    storage = args.storage or ("bf16" if args.config == 2 and not args.adjoint else "fp32")
    project = args.project or ("solve" if args.config in (2, 4) else "reference")
    p = w.params.copy()
    p.storage = STORAGE_NAMES[storage]
    if storage != "fp32" and w.kind != "dcgrid":
        print(json.dumps({"error": "16-bit field storage is implemented for the DCGrid configurations"}))
        sys.exit(2)
    if args.config == 3:
        print(json.dumps(adjoint_bench(w, args)), flush=True)
        return
    if args.adjoint:
        if w.kind != "dcgrid" or storage != "fp32":
            print(json.dumps({"error": "--adjoint times the DCGrid configurations with FP32 storage"}))
            sys.exit(2)
        print(json.dumps(adjoint_dcgrid_bench(w, p, args)), flush=True)
        return
    solve = project == "solve" and w.kind == "dcgrid"
    if solve:
        p.project_mode = PROJECT_SOLVE
        p.solve_precision = SOLVE_MIXED
        p.solve_tol = 1e-8
        p.solve_max_cycles = 50

    log(f"[bench] creating {w.name} {w.size} budget={w.max_num_blocks} storage={storage} project={project}")
    sim = make_sim(dcgrid_amd.Simulation, w, p)
    log(f"[bench] warmup {args.warmup} steps")
    sim.step(args.warmup)
    sim.synchronize()
    act0 = int(sim.info().active_cells)
    sim.set_graph(args.graph and not solve)
    log(f"[bench] timing {args.steps} steps (active cells {act0})")
    cs0 = sim.cell_steps()
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.synchronize()
    t1 = time.perf_counter()
    cell_steps = sim.cell_steps() - cs0  # active cells of every timed step, counted on the device
    info1 = sim.info()
    act1 = int(info1.active_cells)
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = cell_steps / elapsed
    last_solve = sim.solve_report() if solve else None

    # per-kernel pass: every launch of every class bracketed by HIP events on the library's stream
    kernels, roof = {}, None
    if args.timing_steps > 0:
        sim.set_graph(False)
        sim.set_timing(list(dcgrid_amd.K_NAMES))
        cs1 = sim.cell_steps()
        sim.step(args.timing_steps)
        sim.synchronize()
        tcs = sim.cell_steps() - cs1
        tm = sim.timing()
        tb = sim.timing_bytes()
        pc = per_cell_bytes(p)
        for name, (ms, n) in tm.items():
            if n == 0:
                continue
            entry = {"ms_per_step": round(ms / args.timing_steps, 4), "launches_per_step": round(n / args.timing_steps, 1),
                     "avg_launch_us": round(1e3 * ms / n, 3)}
            bytes_ = None
            if name in pc and w.kind == "dcgrid":
                bytes_ = pc[name] * tcs  # every allocated cell once per step
            elif name == "jacobi" and not solve and w.kind == "dcgrid":
                if p.project_top_pairs == p.project_level_pairs:
                    bytes_ = 2 * p.project_level_pairs * 3 * STORAGE_BYTES[p.storage] * tcs
                else:
                    bytes_ = jacobi_bytes(info1, p)[0] * args.timing_steps
            elif name == "jacobi" and w.kind == "uniform" and p.project_mode == 1:
                bytes_ = 12 * int(info1.active_cells) * n
            elif tb.get(name):
                bytes_ = tb[name]
            if bytes_:
                entry["algorithmic_GBps"] = round(bytes_ / (ms * 1e-3) / 1e9, 1)
                entry["algorithmic_bytes_per_launch"] = int(bytes_ / n)
            kernels[name] = entry
        modelled = {k: v for k, v in kernels.items() if "algorithmic_GBps" in v}
        if modelled:
            dom = max(modelled, key=lambda k: modelled[k]["ms_per_step"])
            e = modelled[dom]
            kname = {"solve_smooth": "k_sv_smooth", "jacobi": "k_jacobi" if w.kind == "dcgrid" else "k_u_jacobi",
                     "advect_velocity": "k_advect_velocity", "advect_density": "k_advect_density",
                     "divergence": "k_divergence", "apply_pressure": "k_apply_pressure"}[dom]
            roof = {"kernel": kname, "bound": "hbm", "achieved": e["algorithmic_GBps"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(e["algorithmic_GBps"] / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(w.name, kname), "avg_launch_us": e["avg_launch_us"],
                    "launches": int(e["launches_per_step"] * args.timing_steps), "timed_steps": args.timing_steps,
                    "algorithmic_bytes_per_launch": e["algorithmic_bytes_per_launch"],
                    "share_of_step_kernel_time": round(e["ms_per_step"] / sum(v["ms_per_step"] for v in kernels.values()), 3)}
    solve_cmp = None
    if solve and not args.no_solve_compare:
        # compared on FP32 field storage (16-bit storage rounds the pressures of every variant
        # to the same values), on the same scene after the same number of warm-up steps
        log("[bench] solve comparison (FP32 storage, mixed / FP32 / FP64)")
        sim.close()
        q = p.copy()
        q.storage = STORAGE_NAMES["fp32"]
        sim = make_sim(dcgrid_amd.Simulation, w, q)
        sim.step(args.warmup)
        solve_cmp = solve_compare(sim, q)
        solve_cmp["field_storage"] = "fp32"
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "active cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[p.storage],
        "arithmetic": "f32, f64 residual / norms / update of the refinement" if solve else "f32",
        "data": "synthetic (seeded sphere source + value-noise velocity)" if args.config != 4 else
                "synthetic (seeded fBm terrain, moist layer, ambient wind)",
        "config": {
            "workload": w.name,
            "domain": list(w.size),
            "block_width": p.block_width if w.kind == "dcgrid" else None,
            "levels": int(info1.levels),
            "budget_blocks": w.max_num_blocks if w.kind == "dcgrid" else None,
            "field_storage": storage,
            "projection": "each level solved to 1e-8 relative residual (FP64 refinement around FP32 multigrid), "
                          "every step" if solve else "reference Jacobi schedule",
            "active_cells_first_last": [act0, act1],
            "mean_active_cells": round(cell_steps / args.steps, 1),
            "block_loads": info1.levels_list("block_loads") if w.kind == "dcgrid" else None,
            "launch": "hip graph (pairs of steps)" if (args.graph and not solve) else "eager",
        },
        "roofline": roof,
        "kernels": kernels,
        "last_step_solve": last_solve,
        "solve": solve_cmp,
        "cpu_baseline": None,
    }
    if not solve:
        sb = step_bytes(info1, p, uniform=w.kind == "uniform")
        out["config"]["step_algorithmic_bytes"] = sb
        out["config"]["step_algorithmic_GBps"] = round(sb / (ms_per_step * 1e-3) / 1e9, 1)
    sim.close()
    if not args.no_cpu:
        log(f"[bench] CPU baseline (~{args.cpu_seconds}s)")
        cp = p.copy()
        note = ""
        if solve:
            # the oracle's solve is an exact-to-1e-12 SOR reference (a correctness check, minutes
            # per step at this size): the CPU runs the same scene with the reference projection
            cp.project_mode = 0
            note = "; projection: the reference's Jacobi schedule (the CPU oracle has no multigrid solve)"
        cb = cpu_baseline(w, cp, 2, args.cpu_seconds)
        cb["sample"] = (f"{cb.pop('steps')} steps of {w.name} after 2 untimed steps from the same seeded scene "
                        f"({cb.pop('cells_per_step')} active cells/step), FP32 oracle, {storage} field rounding, "
                        f"OpenMP {cb['cores']} threads{note}")
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
