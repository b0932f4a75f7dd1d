"""Scene runner: the seeded benchmark configurations through the library in the reference's
frame order, with JSON-lines metrics, state checkpoints and resume.

    python -m dcgrid_amd.run --config 1 --steps 200 --every 20 --jsonl run.jsonl
    python -m dcgrid_amd.run --config 4 --steps 300 --checkpoint state.dcg --checkpoint-every 50
    python -m dcgrid_amd.run --config 4 --resume state.dcg --start-step 150 --steps 150

The reference's driver uploads the parameters once per frame (copySimParamsToDevice,
src/utils/sim_utils.h:6, called from src/simulation.cpp:93-98) and then runs advectVelocity,
adaptTopology, project and advectDensity (src/simulation.cpp:100-116); its scenes are
src/data/scenes.h:13-32 over SimParams::defaultParams (src/data/sim_params.cpp:4-34). Here the
parameters go to the library with dcg_set_params before the run (and are part of a saved state),
and one dcg_step is one frame in the same order. The workloads are dcgrid_amd.scenes.config.

Each report line covers `--every` frames: the frames' wall time (HIP-synchronised), active cells
and block loads after the last frame, active cell-steps per second, optionally the reference's
residual (debugStats) and the projection solve report. The first line describes the run, the
last one sums it up. Configuration 3 with --adjoint runs the differentiable rollout instead
(tape forward with checkpoints, adjoint of the L2 distance to the target density) and reports
the loss and gradient norms.
"""
import argparse
import json
import sys
import time

STORAGES = {"fp32": 0, "fp16": 1, "bf16": 2}


def parse_args(argv=None):
    ap = argparse.ArgumentParser(prog="python -m dcgrid_amd.run", description=__doc__.split("\n\n")[0])
    ap.add_argument("--config", type=int, default=1, choices=range(5), help="benchmark configuration")
    ap.add_argument("--scale", type=int, default=1, help="divide the domain by this factor")
    ap.add_argument("--steps", type=int, default=None, help="frames to run (default: the configuration's)")
    ap.add_argument("--every", type=int, default=10, help="frames per report line")
    ap.add_argument("--storage", choices=sorted(STORAGES), default="fp32", help="DCGrid field storage")
    ap.add_argument("--jsonl", default="-", help="metrics output file ('-': standard output)")
    ap.add_argument("--residual", action="store_true", help="add the debugStats residual to each line")
    ap.add_argument("--checkpoint", default=None, help="state file written every --checkpoint-every "
                                                       "frames and at the end")
    ap.add_argument("--checkpoint-every", type=int, default=0)
    ap.add_argument("--resume", default=None, help="state file to continue from")
    ap.add_argument("--start-step", type=int, default=0, help="frame number of the resumed state")
    ap.add_argument("--graph", action="store_true", help="replay frames from a captured HIP graph")
    ap.add_argument("--adjoint", action="store_true", help="configuration 3: differentiable rollout")
    ap.add_argument("--checkpoint-segment", type=int, default=8, help="adjoint checkpoint interval")
    a = ap.parse_args(argv)
    if a.every <= 0:
        ap.error("--every must be positive")
    if a.adjoint and a.config != 3:
        ap.error("--adjoint runs configuration 3")
    return a


class _Out:
    def __init__(self, path):
        self.f = sys.stdout if path == "-" else open(path, "w")

    def line(self, obj):
        self.f.write(json.dumps(obj) + "\n")
        self.f.flush()

    def close(self):
        if self.f is not sys.stdout:
            self.f.close()


def _loads(sim):
    inf = sim.info()
    if sim.kind != "dcgrid":
        return int(inf.active_cells), []
    return int(inf.active_cells), [int(x) for x in list(inf.block_loads)[:inf.levels]]


def _adjoint(args, w, out):
    import torch
    import dcgrid_amd
    from dcgrid_amd import scenes
    from dcgrid_amd.autograd import rollout
    n = w.size[0]
    steps = args.steps or w.steps
    den, vel, tgt = scenes.control_inputs(n)
    dev = torch.device("cuda")
    sim = dcgrid_amd.Simulation(w.kind, w.size, w.params)
    d0 = torch.as_tensor(den, device=dev).requires_grad_(True)
    v0 = torch.as_tensor(vel, device=dev).requires_grad_(True)
    target = torch.as_tensor(tgt, device=dev)
    t0 = time.perf_counter()
    dN, _ = rollout(sim, d0, v0, steps, args.checkpoint_segment)
    loss = 0.5 * ((dN - target) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out.line({"event": "adjoint", "steps": steps, "checkpoint_every": args.checkpoint_segment,
              "loss": float(loss), "grad_norm_density": float(torch.linalg.vector_norm(d0.grad)),
              "grad_norm_velocity": float(torch.linalg.vector_norm(v0.grad)), "wall_s": round(wall, 4)})
    sim.close()


def main(argv=None):
    args = parse_args(argv)
    import dcgrid_amd
    from dcgrid_amd import scenes
    from dcgrid_amd._abi import PROJECT_SOLVE
    w = scenes.config(args.config, scale=args.scale)
    p = w.params.copy()
    if w.kind == "dcgrid":
        p.storage = STORAGES[args.storage]
    elif args.storage != "fp32":
        raise SystemExit("16-bit field storage is implemented for the DCGrid")
    out = _Out(args.jsonl)
    head = {"event": "start", "config": args.config, "workload": w.name, "kind": w.kind, "domain": list(w.size),
            "budget_blocks": w.max_num_blocks if w.kind == "dcgrid" else None, "storage": args.storage,
            "resumed_from": args.resume, "start_step": args.start_step}
    if args.adjoint:
        out.line(dict(head, params=p.as_dict()))
        _adjoint(args, w, out)
        out.close()
        return 0
    sim = dcgrid_amd.Simulation(w.kind, w.size, p, w.max_num_blocks if w.kind == "dcgrid" else None)
    if args.resume:
        sim.load_state(args.resume)  # parameters included
        p = sim.params
    elif args.config == 3:
        den, vel, _ = scenes.control_inputs(w.size[0])
        sim.import_field("density", den)
        sim.import_field("velocity", vel.ravel())
    else:
        sim.seed_scene()
    out.line(dict(head, params=p.as_dict()))
    if args.graph:
        sim.set_graph(True)
    steps = args.steps if args.steps is not None else w.steps
    step = args.start_step
    done = 0
    total_cells = 0
    t_run = time.perf_counter()
    sim.synchronize()
    while done < steps:
        k = min(args.every, steps - done)
        if args.checkpoint and args.checkpoint_every > 0:
            k = min(k, args.checkpoint_every - (step % args.checkpoint_every))
        c0 = sim.cell_steps()
        t0 = time.perf_counter()
        sim.step(k)
        sim.synchronize()
        dt = time.perf_counter() - t0
        cells = sim.cell_steps() - c0
        done += k
        step += k
        total_cells += cells
        active, loads = _loads(sim)
        rec = {"step": step, "frames": k, "ms_per_step": round(1e3 * dt / k, 4), "active_cells": active,
               "active_cell_steps_per_s": round(cells / dt, 1), "sim_time": round(step * float(p.dt), 6)}
        if loads:
            rec["block_loads"] = loads
        if args.residual:
            rec["residual"] = sim.debug_stats()
        if p.project_mode == PROJECT_SOLVE:
            rec["solve"] = sim.solve_report()
        out.line(rec)
        if args.checkpoint and args.checkpoint_every > 0 and step % args.checkpoint_every == 0:
            sim.save_state(args.checkpoint)
    wall = time.perf_counter() - t_run
    if args.checkpoint:
        sim.save_state(args.checkpoint)
    fin = {"event": "end", "steps": done, "last_step": step, "wall_s": round(wall, 4),
           "mean_ms_per_step": round(1e3 * wall / max(done, 1), 4), "active_cell_steps": int(total_cells),
           "checkpoint": args.checkpoint}
    out.line(fin)
    out.close()
    sim.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
