"""Pins of the oracle's 16-bit field storage (dcg_params.storage), the reference the GPU's FP16 /
BF16 storage is checked against bit for bit.

The oracle's rounding restates the formats from their definitions (quantum 2^(E - mantissa bits),
nearest even); it is pinned here against numpy's own binary16 conversion and against an
independent nearest-even construction of bfloat16 from the float's bit pattern. A simulation with
16-bit storage must leave only representable values in every field (a store that skipped the
rounding leaves a value that is not), and stay close to the FP32-storage run."""
import ctypes as C

import numpy as np
import pytest

from dcgrid_amd import scenes
from dcgrid_amd._abi import STORAGE_BF16, STORAGE_FP16, STORAGE_FP32
from parity_util import rel_l2


def round_storage(oracle_lib, x, storage):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    oracle_lib.library().orc_round_storage(x.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                                           C.c_uint64(x.size), storage)
    return out


def edge_values():
    v = [0.0, -0.0, 1.0, -1.0, 65504.0, 65519.99, 65520.0, 65536.0, 1e9, 6.1035156e-05, 6.0e-05,
         5.96e-08, 2.98e-08, 2.9802322e-08, 2.9802326e-08, 1e-10, 3.4e38, -3.4e38, 1.1754944e-38, 1e-40,
         np.inf, -np.inf]
    # exact ties of both formats around 1 and in the binary16 subnormal range
    v += [1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 2.0 ** -25, 3 * 2.0 ** -25]
    return np.array(v + [-a for a in v], dtype=np.float32)


def random_floats(n, seed):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    # plus a dense sweep over the binary16 range (normal, subnormal, overflow boundary)
    sweep = np.arange(0x33000000, 0x47900000, 4099, dtype=np.uint32).view(np.float32)
    return np.concatenate([x, sweep, -sweep, edge_values()])


def test_fp16_rounding_matches_numpy(oracle_lib):
    x = random_floats(1_500_000, 1)
    got = round_storage(oracle_lib, x, STORAGE_FP16)
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def bf16_nearest_even(x):
    """Nearest bfloat16 by exact distance to the two neighbouring bfloat16 values (float64
    arithmetic), ties to the even one."""
    u = x.view(np.uint32).astype(np.uint64)
    lo = (u & 0xFFFF0000).astype(np.uint32)
    hi = (lo.astype(np.uint64) + 0x10000).astype(np.uint32)
    flo, fhi = lo.view(np.float32).astype(np.float64), hi.view(np.float32).astype(np.float64)
    xd = x.astype(np.float64)
    dlo, dhi = np.abs(xd - flo), np.abs(fhi - xd)
    take_hi = (dhi < dlo) | ((dhi == dlo) & (((lo >> 16) & 1) == 1))
    return np.where(take_hi, hi, lo).view(np.float32)


def test_bf16_rounding_matches_nearest_even(oracle_lib):
    x = random_floats(1_500_000, 2)
    x = x[np.isfinite(x) & (np.abs(x) < 3.38e38)]  # the construction needs a finite upper neighbour
    got = round_storage(oracle_lib, x, STORAGE_BF16)
    ref = bf16_nearest_even(x)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_fp32_storage_is_identity(oracle_lib):
    x = random_floats(100_000, 3)
    assert np.array_equal(round_storage(oracle_lib, x, STORAGE_FP32).view(np.uint32), x.view(np.uint32))


def run(oracle_lib, storage, steps=4, name="cfg1_64"):
    w = scenes.config(1, scale=4) if name == "cfg1_64" else scenes.config(4, scale=8)
    p = w.params.copy()
    p.storage = storage
    s = oracle_lib.OracleSim(w.kind, w.size, p, w.max_num_blocks, precision=32)
    s.seed_scene()
    s.step(steps)
    return s


FIELDS = ("density", "velocity", "temperature", "pressure", "t_pressure", "divergence", "fluidity", "vorticity")


@pytest.mark.parametrize("storage", [STORAGE_FP16, STORAGE_BF16])
@pytest.mark.parametrize("name", ["cfg1_64", "cfg4_64"])
def test_every_stored_value_is_representable(oracle_lib, storage, name):
    s = run(oracle_lib, storage, name=name)
    inf = s.info()
    m = np.repeat(s.export_tables()["block_levels"] != 0xFF, inf.block_width ** 3)
    fields = FIELDS + (("vapor",) if s.params.moist else ())
    for f in fields:
        v = s.export_field(f)
        v = v.reshape(-1, 3)[m].ravel() if v.size == 3 * m.size else v[m]
        assert np.array_equal(round_storage(oracle_lib, v, storage).view(np.uint32), v.view(np.uint32)), f
    s.close()


def test_storage_runs_stay_close_to_fp32(oracle_lib):
    ref = run(oracle_lib, STORAGE_FP32)
    m = np.repeat(ref.export_tables()["block_levels"] != 0xFF, ref.info().block_width ** 3)
    # measured after 4 steps: FP16 1e-4..7e-4, BF16 1e-3..5e-3 (relative L2)
    for storage, tol in ((STORAGE_FP16, 2e-3), (STORAGE_BF16, 2e-2)):
        s = run(oracle_lib, storage)
        # the rounding moves no block in this scene: the fields are comparable cell by cell
        same = all(np.array_equal(s.export_tables()[k], ref.export_tables()[k])
                   for k in ("block_levels", "block_positions"))
        assert same
        for f in ("density", "velocity", "pressure", "temperature"):
            a, b = s.export_field(f), ref.export_field(f)
            if f == "velocity":
                a, b = a.reshape(-1, 3)[m], b.reshape(-1, 3)[m]
            else:
                a, b = a[m], b[m]
            assert rel_l2(a, b) <= tol, (storage, f, rel_l2(a, b))
        s.close()
    ref.close()
