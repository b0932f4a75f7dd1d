"""Pins of the terrain / moist-atmosphere extension (no reference code): the deterministic exp
against numpy, the Magnus saturation curve, the saturation adjustment's conservation laws, the
terrain heightfield and its fluidity, and the seeded lapse-rate state."""
import ctypes as C

import numpy as np
import pytest

from dcgrid_amd import scenes


@pytest.fixture(scope="module")
def lib(oracle_lib):
    return oracle_lib.library()


def params():
    return scenes.config(4, scale=8).params


def test_exp_det_matches_numpy(lib):
    xs = np.linspace(-20.0, 20.0, 4001, dtype=np.float32)
    got = np.array([lib.orc_exp_det(float(x)) for x in xs], dtype=np.float64)
    ref = np.exp(xs.astype(np.float64))
    assert np.max(np.abs(got - ref) / ref) <= 4e-7
    assert lib.orc_exp_det(0.0) == 1.0


def test_magnus_saturation_curve(lib):
    p = params()
    pp = C.byref(p)
    assert lib.orc_qsat(pp, 0.0) == C.c_float(p.qsat_scale).value
    for t in (-30.0, -5.0, 10.0, 20.0, 35.0):
        ref = p.qsat_scale * np.exp(17.27 * t / (t + 237.3))
        assert abs(lib.orc_qsat(pp, t) - ref) <= 1e-6 * ref
    # monotone in temperature
    ts = np.linspace(-40, 40, 81)
    q = [lib.orc_qsat(pp, float(t)) for t in ts]
    assert np.all(np.diff(q) > 0)


def adjust(lib, p, qv, qc, t):
    a, b, c = C.c_float(qv), C.c_float(qc), C.c_float(t)
    lib.orc_saturation_adjust(C.byref(p), C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def test_saturation_adjustment_conserves_water_and_heats(lib):
    p = params()
    rng = np.random.Generator(np.random.PCG64(9))
    for _ in range(200):
        t = float(rng.uniform(-20, 30))
        qs = lib.orc_qsat(C.byref(p), t)
        qv = float(qs * rng.uniform(0.5, 1.5))
        qc = float(rng.uniform(0.0, 2e-3))
        qv1, qc1, t1 = adjust(lib, p, qv, qc, t)
        # total water is conserved to float rounding; cloud water never negative
        assert abs((qv1 + qc1) - (qv + qc)) <= 4 * np.spacing(np.float32(qv + qc))
        assert qc1 >= 0.0
        c = qc1 - qc
        # latent heating proportional to the condensed amount
        assert abs((t1 - t) - p.latent_heat * c) <= 1e-4 * max(1.0, abs(t1 - t)) + 4 * np.spacing(np.float32(t1))
        if qv > qs:
            assert c > 0 and abs(c - p.cond_rate * (qv - qs)) <= 1e-6 * qs
        elif qc > 0:
            assert c <= 0 and c >= -qc


def test_terrain_height_range_and_fluidity(lib):
    p = params()
    pp = C.byref(p)
    hs = np.array([[lib.orc_terrain_height(pp, float(x), float(z)) for x in range(0, 128, 4)] for z in range(0, 128, 4)])
    assert hs.min() >= 0.0 and hs.max() <= p.terrain_height
    assert hs.std() > 0.05 * p.terrain_height  # not flat
    # a cell well below the terrain is solid, one well above is fluid, at every level
    for x, z in ((10, 20), (70, 90)):
        h = lib.orc_terrain_height(pp, x + 0.5, z + 0.5)
        below, above = int(h) - 3, int(h) + 3
        if below >= 0:
            assert lib.orc_cell_fluidity(pp, x, below, z, 1) == 0.0
        assert lib.orc_cell_fluidity(pp, x, above, z, 1) == 1.0


def test_seeded_lapse_rate_state(oracle_lib):
    w = scenes.config(4, scale=8)
    p = w.params
    s = oracle_lib.OracleSim(w.kind, w.size, p, w.max_num_blocks, precision=64)
    s.seed_scene()
    inf = s.info()
    t = s.export_tables()
    lv = t["block_levels"]
    pos = t["block_positions"].reshape(-1, 3)
    tem = s.export_field_f64("temperature").reshape(-1, 512)
    vap = s.export_field_f64("vapor").reshape(-1, 512)
    # level-0 block 0 column: T = t_ambient - lapse * height at every cell centre
    b = int(np.flatnonzero(lv == 0)[0])
    from oracle import library
    lib_ = library()
    for loc in range(512):
        y = ((loc >> 1) & 1) | (((loc >> 4) & 1) << 1) | (((loc >> 7) & 1) << 2)
        hy = np.float32(pos[b, 1] + y) + np.float32(0.5)
        ta = np.float32(p.t_ambient) - np.float32(p.lapse_rate) * hy
        assert tem[b, loc] == ta
        qa = np.float32(p.humidity) * np.float32(lib_.orc_qsat(C.byref(p), float(ta))) if hy < p.moist_layer else 0.0
        assert vap[b, loc] == np.float32(qa)
    s.close()
    assert inf.levels == p.max_levels
