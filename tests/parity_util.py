"""Shared helpers of the parity tests: run the HIP library and the CPU oracle side by side."""
import numpy as np

TABLE_KEYS = ("block_positions", "block_levels", "block_flags", "parent_indices", "child_indices", "cell_indices")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def active_cell_mask(sim):
    """Boolean mask of the cells of allocated blocks (DCGrid) or all cells (uniform)."""
    inf = sim.info()
    if inf.kind == 0:
        return None
    t = sim.export_tables()
    bw3 = int(inf.block_width) ** 3
    return np.repeat(t["block_levels"] != 0xFF, bw3)


def field_on(sim, name, mask):
    f = sim.export_field(name)
    if mask is not None:
        f = f[mask]
    return f


def compare_tables(ta, tb, keys=TABLE_KEYS):
    """Names of the tables that differ (bit for bit)."""
    return [k for k in keys if not np.array_equal(ta[k], tb[k])]


def make_pair(kind, size, params, max_blocks=None, precision=64):
    import dcgrid_amd
    import oracle
    g = dcgrid_amd.Simulation(kind, size, params, max_blocks)
    c = oracle.OracleSim(kind, size, params, max_blocks, precision=precision)
    return g, c
