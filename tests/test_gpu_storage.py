"""16-bit field storage on the GPU (dcg_params.storage = FP16 / BF16) through the C ABI.

Bar: the GPU run is bit-exact against the oracle's FP32 instantiation with the same storage
(every store rounded to nearest even, FP32 arithmetic), tables included; import rounds like
numpy's binary16 conversion and export widens exactly; the projection solve still reaches its
tolerance (the iterate is FP64) and stores the rounded solution; the 16-bit runs stay within a
stated distance of the FP32-storage run."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dcgrid_amd  # noqa: E402
import oracle  # noqa: E402
from dcgrid_amd import scenes  # noqa: E402
from dcgrid_amd._abi import STORAGE_BF16, STORAGE_FP16, STORAGE_FP32, SOLVE_MIXED  # noqa: E402
from parity_util import TABLE_KEYS, compare_tables, rel_l2  # noqa: E402

FIELDS = ("density", "velocity", "temperature", "pressure", "t_pressure", "divergence", "fluidity", "vorticity")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def workload(name, storage):
    if name == "ref64_bw4":
        p = scenes.reference_params(64)
        kind, size, b, seed = "dcgrid", (64, 64, 64), 1500, False
    elif name == "cfg1_64":
        w = scenes.config(1, scale=4)
        kind, size, p, b, seed = w.kind, w.size, w.params, w.max_num_blocks, True
    elif name == "cfg4_128":
        w = scenes.config(4, scale=8)
        kind, size, p, b, seed = w.kind, w.size, w.params, w.max_num_blocks, True
    else:
        raise KeyError(name)
    p = p.copy()
    p.storage = storage
    return kind, size, p, b, seed


def allocated(sim):
    return np.repeat(sim.export_tables()["block_levels"] != 0xFF, sim.info().block_width ** 3)


def bits_equal(a, b, m):
    if a.size == 3 * m.size:
        a, b = a.reshape(-1, 3)[m], b.reshape(-1, 3)[m]
    else:
        a, b = a[m], b[m]
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("storage", [STORAGE_FP16, STORAGE_BF16])
@pytest.mark.parametrize("name", ["ref64_bw4", "cfg1_64", "cfg4_128"])
def test_steps_bitwise_vs_oracle_with_same_storage(name, storage):
    kind, size, p, b, seed = workload(name, storage)
    say(f"[case] {name} storage={storage}")
    g = dcgrid_amd.Simulation(kind, size, p, b)
    c = oracle.OracleSim(kind, size, p, b, precision=32)
    if seed:
        g.seed_scene()
        c.seed_scene()
    for step in range(4):
        g.step(1)
        c.step(1)
        assert compare_tables(g.export_tables(), c.export_tables(), TABLE_KEYS) == [], (name, step)
        m = allocated(c)
        for f in FIELDS + (("vapor",) if p.moist else ()):
            assert bits_equal(g.export_field(f), c.export_field(f), m), (name, storage, step, f)
    g.close()
    c.close()


@pytest.mark.parametrize("storage", [STORAGE_FP16, STORAGE_BF16])
def test_import_rounds_and_export_widens(storage):
    kind, size, p, b, _ = workload("ref64_bw4", storage)
    g = dcgrid_amd.Simulation(kind, size, p, b)
    n = g.field_size("density")
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(n) * 10.0 ** rng.uniform(-9, 4, n)).astype(np.float32)
    x[:4] = [65519.0, 65520.0, 1e-8, -3e-6]
    g.import_field("density", x)
    y = g.export_field("density")
    ref = np.empty_like(x)
    oracle.library().orc_round_storage(x.ctypes.data, ref.ctypes.data, n, storage)
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))
    if storage == STORAGE_FP16:
        with np.errstate(over="ignore"):
            assert np.array_equal(y.view(np.uint32), x.astype(np.float16).astype(np.float32).view(np.uint32))
    g.close()


@pytest.mark.parametrize("storage", [STORAGE_FP16, STORAGE_BF16])
def test_projection_solve_with_16bit_storage(storage):
    """The solve's iterate is FP64 and its divergence input the stored one: it converges as with
    FP32 storage, and the stored pressure is the oracle's solution rounded once."""
    kind, size, p, b, seed = workload("cfg1_64", storage)
    p.solve_precision = SOLVE_MIXED
    g = dcgrid_amd.Simulation(kind, size, p, b)
    c = oracle.OracleSim(kind, size, p, b, precision=64)
    g.seed_scene()
    c.seed_scene()
    g.step(3)
    c.step(3)
    c.import_field("velocity", g.export_field("velocity"))
    g.project_solve()
    c.project_solve()
    rep = g.solve_report()
    say("report", rep)
    assert rep["converged"]
    m = allocated(g)
    pg, pc = g.export_field("pressure")[m], c.export_field_f64("pressure")[m]
    # one rounding of the FP64 solution: relative error at most half a unit in the last place
    ulp = 2.0 ** -11 if storage == STORAGE_FP16 else 2.0 ** -8
    err = rel_l2(pg, pc)
    say("pressure rel l2", err)
    assert err <= ulp
    g.close()
    c.close()


def test_16bit_runs_close_to_fp32():
    out = {}
    for storage in (STORAGE_FP32, STORAGE_FP16, STORAGE_BF16):
        kind, size, p, b, seed = workload("cfg1_64", storage)
        g = dcgrid_amd.Simulation(kind, size, p, b)
        g.seed_scene()
        g.step(4)
        out[storage] = (g.export_tables()["block_levels"], {f: g.export_field(f) for f in ("density", "velocity")})
        g.close()
    m = np.repeat(out[STORAGE_FP32][0] != 0xFF, 512)
    for storage, tol in ((STORAGE_FP16, 5e-3), (STORAGE_BF16, 5e-2)):
        assert np.array_equal(out[storage][0], out[STORAGE_FP32][0])
        for f in ("density", "velocity"):
            a, b_ = out[storage][1][f], out[STORAGE_FP32][1][f]
            a, b_ = (a.reshape(-1, 3)[m], b_.reshape(-1, 3)[m]) if f == "velocity" else (a[m], b_[m])
            e = rel_l2(a, b_)
            say("storage", storage, f, "rel l2 vs fp32", e)
            assert e <= tol
