"""Workloads of the DCGrid adjoint tests: seeded scenes advanced a few steps with adaptation
(moves and refinement), so the rollout's frozen block topology has sparse, refined levels and
cross-level samples, divergence and gradient stencils.

  cfg1_64           configuration 1 at 64^3: 8^3 blocks, 3 levels, a 64-block budget (levels 0
                    and 1 sparse), sphere source with temperature and buoyancy, multilevel
                    Jacobi projection (5 + 5 pairs)
  cfg1_64_local     the same with the local projection (4 pairs per level)
  ref64_bw4_solids  the reference defaults at 64^3: 4^3 blocks, 5 levels, the inlet, additional
                    solids (fluidity-weighted sampling, divergence and gradient)
"""
from dcgrid_amd import scenes
from dcgrid_amd._abi import PROJECT_LOCAL

NAMES = ["cfg1_64", "cfg1_64_local", "ref64_bw4_solids"]


def workload(name):
    """(kind, size, params, block budget, adaptive steps before the rollout)."""
    if name in ("cfg1_64", "cfg1_64_local"):
        w = scenes.config(1, scale=4)
        p = w.params
        if name == "cfg1_64_local":
            p.project_mode = PROJECT_LOCAL
            p.project_local_pairs = 4
        return w.kind, w.size, p, w.max_num_blocks, 4
    if name == "ref64_bw4_solids":
        p = scenes.reference_params(64)
        p.enable_additional_solids = 1
        return "dcgrid", (64, 64, 64), p, 2000, 3
    raise KeyError(name)


def prepared(make, name):
    """A simulation of `name` (make(kind, size, params, budget)), seeded and advanced by its
    adaptive steps."""
    kind, size, p, b, pre = workload(name)
    s = make(kind, size, p, b)
    s.seed_scene()
    s.step(pre)
    return s


def allocated(sim):
    """Cells of allocated blocks."""
    import numpy as np
    inf = sim.info()
    return np.repeat(sim.export_tables()["block_levels"] != 0xFF, inf.block_width ** 3)
