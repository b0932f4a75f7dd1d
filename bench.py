#!/usr/bin/env python3
"""Benchmark of the DCGrid simulation step on one MI355X.

Workload (default): the adaptive smoke plume on a 256^3 virtual domain of 8^3-cell blocks, 3
levels, a 4096-block budget, feature-driven refinement (density gradient 0.05, vorticity 0.5),
the buoyant sphere source scaled 4x from the 64^3 plume, and the seeded 3-octave value-noise
initial velocity of amplitude 0.2. Block tables are rebuilt every step. Synthetic, seeded inputs.

One step = advectVelocity -> adaptTopology -> project -> advectDensity (src/simulation.cpp:104-111)
through the C ABI. Prints one JSON line (see README of the contract in DESIGN.md):
  value          active cell-steps/s (sum of allocated cells over the timed steps / time)
  roofline       the Jacobi half-sweep kernel (the projection's hot loop): algorithmic bytes
                 (12 B per cell per sweep: read p and div, write p~) over its HIP-event-timed
                 average duration, against the 8 TB/s HBM peak
  cpu_baseline   the CPU oracle (FP32 restatement, OpenMP over all host cores) on a bounded
                 sample of the same workload
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


STORAGE_BYTES = {0: 4, 1: 2, 2: 2}  # dcg_params.storage -> bytes per stored scalar
STORAGE_NAMES = {"fp32": 0, "fp16": 1, "bf16": 2}


def jacobi_bytes(info, params):
    """Algorithmic bytes and launch count of the Jacobi half-sweeps of one project() call: each
    half-sweep reads p and div and writes p (three stored scalars per cell)."""
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


def step_bytes(info, params):
    """Algorithmic HBM bytes of one step (per-stage compulsory field traffic, no solids); S is the
    stored scalar size (4 for FP32 fields, 2 for FP16/BF16), V = 3 S."""
    bv = int(info.block_width) ** 3
    L = int(info.levels)
    S = STORAGE_BYTES[params.storage]
    V = 3 * S
    src = params.source_mode != 0
    k = 2 if src else 1  # advected scalars
    act = sum(int(info.block_loads[l]) for l in range(L)) * bv
    below = sum(int(info.block_loads[l]) for l in range(L - 1)) * bv
    jb, _ = jacobi_bytes(info, params)
    b = act * (2 * V + (2 * S if src else 0))          # advect u (+ buoyancy)
    b += below * 9 * V // 8 * 2                         # restrict u twice
    b += act * (V + S) + below * 9 * S // 8             # divergence + restrict div
    b += jb                                             # Jacobi sweeps
    b += sum(int(info.block_loads[l]) for l in range(L - 1)) * bv * S * 9 // 8  # prolongate
    b += act * (2 * V + S)                              # gradient subtraction
    b += act * (V + 2 * k * S) + below * 9 * k * S // 8  # advect scalars + restrict
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


def make_sim(cls, w, **kw):
    s = cls(w.kind, w.size, w.params, w.max_num_blocks if w.kind == "dcgrid" else None, **kw)
    s.seed_scene()
    return s


def cpu_baseline(w, warm_steps, budget_s):
    """The oracle (FP32) on the same seeded workload: warm_steps untimed, then steps until
    about budget_s seconds of CPU work; reports active cell-steps/s."""
    import oracle
    oracle.library()
    cores = oracle.get_threads()
    c = make_sim(oracle.OracleSim, w, precision=32)
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
        "sample": f"{n} steps of {w.name} after {warm_steps} untimed steps from the same seeded scene "
                  f"({cells / max(n, 1):.0f} active cells/step), FP32 oracle, OpenMP {cores} threads",
        "seconds": round(dt, 3),
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
        "forward_cell_steps_per_s": round(N * steps / t_fwd, 1),
        "grad_norm_velocity": float(torch.linalg.vector_norm(ov)),
    }


def solve_bench(w, args):
    """Projection solve of the adaptive workload: after `warmup` reference steps, the same
    velocity field is projected with every level solved to a 1e-8 relative residual by FP64
    iterative refinement around FP32 multigrid (mixed), and by the all-FP32 and all-FP64
    variants; wall time of each projection (median of `steps` repetitions from the same
    state), cycles and reached residual."""
    import dcgrid_amd
    from dcgrid_amd._abi import SOLVE_FP32, SOLVE_FP64, SOLVE_MIXED
    p = w.params.copy()
    p.solve_tol = 1e-8
    p.solve_max_cycles = 60
    sim = make_sim(dcgrid_amd.Simulation, w)
    sim.set_params(p)
    log(f"[bench] solve: {args.warmup} reference steps")
    sim.step(args.warmup)
    vel = sim.export_field("velocity")
    inf = sim.info()
    out = {"metric": "projection solve time to 1e-8 relative residual", "unit": "ms",
           "higher_is_better": False, "n_gpus": 1, "data": "synthetic (seeded plume after warmup steps)",
           "config": {"workload": w.name, "domain": list(w.size), "block_loads": inf.levels_list("block_loads"),
                      "active_cells": int(inf.active_cells)}}
    reps = max(1, min(args.steps, 5))
    for name, prec in (("mixed", SOLVE_MIXED), ("fp32", SOLVE_FP32), ("fp64", SOLVE_FP64)):
        q = p.copy()
        q.solve_precision = prec
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
        times.sort()
        out[name] = {"ms": round(times[len(times) // 2], 3), "cycles": rep["cycles"],
                     "rel_residual": [float(f"{r:.3e}") for r in rep["rel_residual"]],
                     "converged": rep["converged"]}
        log(f"[bench] {name}: {out[name]}")
    out["value"] = out["mixed"]["ms"]
    sim.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # default: the configuration's full 200-step run (20 untimed + 180 timed)
    ap.add_argument("--steps", type=int, default=180)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="skip the per-kernel event timing pass")
    ap.add_argument("--timing-steps", type=int, default=20,
                    help="steps of the separate pass that times every Jacobi launch with HIP events")
    ap.add_argument("--storage", choices=sorted(STORAGE_NAMES), default="fp32",
                    help="field storage format in HBM (FP32 arithmetic either way)")
    ap.add_argument("--graph", action="store_true",
                    help="replay the timed steps in pairs from one captured HIP graph (measured no "
                         "faster than eager launches on MI355X for this step)")
    ap.add_argument("--solve", action="store_true",
                    help="time the projection solve to 1e-8 (mixed, FP32, FP64) instead of the step")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 1 GiB device-to-device copy first (calibrates rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(the sharding design is described in DESIGN.md)"}))
        sys.exit(2)

    import dcgrid_amd
    from dcgrid_amd import scenes
    w = scenes.config(args.config)
    w.params.storage = STORAGE_NAMES[args.storage]
    if args.storage != "fp32" and w.kind != "dcgrid":
        print(json.dumps({"error": "16-bit field storage is implemented for the DCGrid configurations"}))
        sys.exit(2)
    if args.config == 3:
        print(json.dumps(adjoint_bench(w, args)), flush=True)
        return
    if args.solve:
        print(json.dumps(solve_bench(w, args)), flush=True)
        return

    if args.copy_probe:
        copy_probe()
    log(f"[bench] creating {w.name} {w.size} budget={w.max_num_blocks}")
    sim = make_sim(dcgrid_amd.Simulation, w)
    log(f"[bench] warmup {args.warmup} steps")
    sim.step(args.warmup)
    sim.synchronize()
    info0 = sim.info()
    act0 = int(info0.active_cells)
    # timed steps: whole-step throughput (per-kernel event timing runs in a pass of its own)
    sim.set_graph(args.graph)
    log(f"[bench] timing {args.steps} steps (active cells {act0}, graph {args.graph})")
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

    roof = None
    if not args.no_timing and args.timing_steps > 0:
        # roofline pass: the following steps with every Jacobi launch bracketed by HIP events on
        # the simulation's stream
        sim.set_graph(False)
        sim.set_timing(["jacobi"])
        cs1 = sim.cell_steps()
        sim.step(args.timing_steps)
        sim.synchronize()
        tcs = sim.cell_steps() - cs1
        info2 = sim.info()
        tm = sim.timing()
        j_ms, j_n = tm["jacobi"]
        p = w.params
        if p.project_top_pairs == p.project_level_pairs:
            # every active cell is swept 2*pairs times per step: exact over the varying topology
            jb = 2 * p.project_level_pairs * 3 * STORAGE_BYTES[p.storage] * tcs
            jl = j_n
        else:
            jb, jl = jacobi_bytes(info2, p)
        if j_n > 0 and j_ms > 0:
            per_launch_bytes = jb / jl
            avg_ms = j_ms / j_n
            achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
            roof = {
                "kernel": "k_jacobi (all levels)",
                "bound": "hbm",
                "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(w.name, "k_jacobi"),
                "avg_launch_us": round(avg_ms * 1e3, 3),
                "launches": int(j_n),
                "timed_steps": args.timing_steps,
                "algorithmic_bytes_per_launch": int(per_launch_bytes),
            }
    sb = step_bytes(info1, w.params)
    out = {
        "metric": "active cell-steps/s",
        "value": round(value, 1),
        "unit": "active cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded sphere source + value-noise velocity)",
        "config": {
            "workload": w.name,
            "domain": list(w.size),
            "block_width": w.params.block_width,
            "levels": int(info1.levels),
            "budget_blocks": w.max_num_blocks,
            "active_cells_first_last": [act0, act1],
            "mean_active_cells": round(cell_steps / args.steps, 1),
            "block_loads": info1.levels_list("block_loads"),
            "step_algorithmic_bytes": sb,
            "step_algorithmic_GBps": round(sb / (ms_per_step * 1e-3) / 1e9, 1),
            "launch": "hip graph (pairs of steps)" if args.graph else "eager",
            "field_storage": args.storage,
        },
        "roofline": roof,
        "cpu_baseline": None,
    }
    sim.close()
    if not args.no_cpu:
        log(f"[bench] CPU baseline (~{args.cpu_seconds}s)")
        out["cpu_baseline"] = cpu_baseline(w, 2, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
