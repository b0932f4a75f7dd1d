"""Pins of the CPU oracle against facts fixed independently of its code: known-answer tests of
the level and budget sizing, textbook Morton order, brute-force geometry on tiny grids, closed
forms of the stencils and transfer operators, and structural invariants of the block tables. CPU
only.
"""
import numpy as np
import pytest

import oracle
from dcgrid_amd import scenes
from dcgrid_amd._abi import (DCGError, DCG_E_ARG, DCG_E_BUDGET, DCG_E_UNSUPPORTED, STAGE_APPLY_PRESSURE,
                             STAGE_DIVERGENCE, STAGE_JACOBI, STAGE_PROLONGATE, STAGE_RESTRICT)

NF = np.uint64(0xFFFFFFFFFFFFFFFF)


# ----------------------------------------------------------------------------- planner KATs
# Independent sizing table (fluid_simulation_dcgrid.cu:29-69 applied by hand).
@pytest.mark.parametrize("size,bw,maxlev,budget,expect_max,expect_sparse", [
    ((256, 256, 256), 8, 3, 4096, [1792, 1792, 512], 2),
    ((512, 512, 512), 8, 4, 32768, [14080, 14080, 4096, 512], 2),
    ((1024, 2048, 1024), 4, 0, 1750000, [575407, 575407, 524288, 65536, 8192, 1024, 128, 16, 2], 2),
    ((1024, 256, 1024), 8, 4, 599040, [524288, 65536, 8192, 1024], 0),
])
def test_planner_kat(size, bw, maxlev, budget, expect_max, expect_sparse):
    p = scenes.reference_params(*size)
    p.block_width = bw
    p.max_levels = maxlev
    inf = oracle.plan(size, budget, p)
    assert inf.levels_list("max_blocks_level") == expect_max
    assert inf.sparse_levels == expect_sparse
    offs = inf.levels_list("level_offsets")
    assert offs == [int(v) for v in np.concatenate([[0], np.cumsum(expect_max)[:-1]])]


def test_planner_levels_reference_rule():
    # levels = 1 + #{k : 2^k divides all dims and 2^k * BW <= minDim} (fluid_simulation_dcgrid.cu:12-22)
    for size, bw, expect in [((256, 256, 256), 4, 7), ((256, 256, 256), 8, 6), ((64, 64, 64), 4, 5),
                             ((96, 64, 64), 4, 5)]:
        p = scenes.reference_params(*size)
        p.block_width = bw
        p.max_levels = 0
        big = 10 ** 7
        if size == (96, 64, 64):
            with pytest.raises(DCGError) as e:  # coarsest blocks would be partial: rejected
                oracle.plan(size, big, p)
            assert e.value.code == DCG_E_UNSUPPORTED
            continue
        assert oracle.plan(size, big, p).levels == expect


def test_planner_errors():
    p = scenes.reference_params(256)
    p.block_width = 4
    p.max_levels = 3
    with pytest.raises(DCGError) as e:  # 3 levels of 4^3 blocks at 256^3 need more than 4096 blocks
        oracle.plan((256, 256, 256), 4096, p)
    assert e.value.code == DCG_E_BUDGET
    p.block_width = 8
    with pytest.raises(DCGError) as e:
        oracle.plan((256, 256, 256), 100, p)  # fewer than the 512 coarsest blocks
    assert e.value.code == DCG_E_BUDGET
    p.block_width = 5
    with pytest.raises(DCGError) as e:
        oracle.plan((256, 256, 256), 4096, p)
    assert e.value.code == DCG_E_UNSUPPORTED


# ----------------------------------------------------------------------------- index algebra
def _textbook_spread(v, bits=11):
    s = 0
    for i in range(bits):
        s |= ((v >> i) & 1) << (3 * i)
    return s


def test_bit_spread_and_hash_textbook():
    lib = oracle.library()
    for v in range(2048):
        assert lib.orc_bit_spread3(v) == _textbook_spread(v)
    rng = np.random.default_rng(0)
    for bw in (4, 8):
        for _ in range(500):
            x, y, z = (int(t) * bw for t in rng.integers(0, 2048, 3))
            want = (_textbook_spread(x // bw) << 2) | (_textbook_spread(y // bw) << 1) | _textbook_spread(z // bw)
            assert lib.orc_grid_hash(x, y, z, bw) == want & 0xFFFFFFFF


def test_local_index_layout():
    lib = oracle.library()
    # reference BW=4 layout [xh yh zh xl yl zl] (dcgrid_utils.cuh:50-58,69-81)
    for x in range(4):
        for y in range(4):
            for z in range(4):
                want = 32 * (x >> 1) + 16 * (y >> 1) + 8 * (z >> 1) + 4 * (x & 1) + 2 * (y & 1) + (z & 1)
                assert lib.orc_local_index(x, y, z, 4) == want
    seen = set()
    for x in range(8):
        for y in range(8):
            for z in range(8):
                i = lib.orc_local_index(x, y, z, 8)
                # top triple = refinement subblock (4^3 cells), lowest triple = restriction group
                assert i >> 6 == 4 * (x >> 2) + 2 * (y >> 2) + (z >> 2)
                assert i & 7 == 4 * (x & 1) + 2 * (y & 1) + (z & 1)
                seen.add(i)
    assert seen == set(range(512))


def test_mipmap_layout_bruteforce():
    lib = oracle.library()
    w, h, d = 16, 8, 12
    total = lib.orc_mipmap_cells(w, h, d)
    assert total == 16 * 8 * 12 + 8 * 4 * 6 + 4 * 2 * 3  # s=1,2,4 divide all dims; s=8 does not divide 12
    seen = []
    off = 0
    for s in (1, 2, 4):
        W, H, D = w // s, h // s, d // s
        for z in range(D):
            for y in range(H):
                for x in range(W):
                    i = lib.orc_mipmap_idx(x, y, z, w, h, d, s)
                    assert i == off + (z * H + y) * W + x
                    seen.append(i)
        off += W * H * D
    assert sorted(seen) == list(range(total))


# ----------------------------------------------------------------------------- geometry helpers
def owners(tables, inf, start_level):
    """Volume at level-`start_level` resolution holding, per cell, the finest allocated block at a
    level >= start_level whose region contains it (-1 if none). Brute force by painting."""
    bw = inf.block_width
    gx, gy, gz = (int(s) >> start_level for s in inf.size)
    vol = -np.ones((gx, gy, gz), np.int64)
    lev = tables["block_levels"]
    pos = tables["block_positions"]
    for level in range(inf.levels - 1, start_level - 1, -1):
        sh = level - start_level
        for b in np.nonzero(lev == level)[0]:
            x, y, z = (int(v) << sh for v in pos[b])
            e = bw << sh
            vol[x:x + e, y:y + e, z:z + e] = b
    return vol


def local_index(x, y, z, bw):
    return oracle.library().orc_local_index(int(x), int(y), int(z), bw)


def make_dcgrid(kind_steps, precision=64):
    name, steps = kind_steps
    if name == "ref64_bw4":
        p = scenes.reference_params(64)
        s = oracle.OracleSim("dcgrid", (64, 64, 64), p, 1500, precision=precision)
    elif name == "ref32_bw8":
        p = scenes.reference_params(64)
        p.block_width = 8
        s = oracle.OracleSim("dcgrid", (64, 64, 64), p, 300, precision=precision)
    else:
        w = scenes.config(1, scale=4)
        s = oracle.OracleSim(w.kind, w.size, w.params, w.max_num_blocks, precision=precision)
        s.seed_scene()
    s.step(steps)
    return s


SCENARIOS = [("ref64_bw4", 0), ("ref64_bw4", 2), ("ref32_bw8", 1), ("cfg1_64", 4)]


@pytest.mark.parametrize("scenario", SCENARIOS)
def test_lookup_bruteforce(scenario):
    s = make_dcgrid(scenario)
    inf = s.info()
    t = s.export_tables()
    vol = owners(t, inf, 0)
    rng = np.random.default_rng(1)
    pts = rng.integers(0, inf.size[0], size=(1500, 3))
    for x, y, z in pts:
        assert s.lookup(int(x), int(y), int(z), 0) == vol[x, y, z]


@pytest.mark.parametrize("scenario", SCENARIOS)
def test_tables_invariants(scenario):
    s = make_dcgrid(scenario)
    inf = s.info()
    t = s.export_tables()
    bw, L = inf.block_width, inf.levels
    lev, pos, par, chi = t["block_levels"], t["block_positions"], t["parent_indices"], t["child_indices"]
    loads = inf.levels_list("block_loads")
    offs = inf.levels_list("level_offsets")
    for level in range(L):
        idx = np.nonzero(lev == level)[0]
        # allocated blocks of a level are a prefix of its slice (free list is iota, no deletion)
        assert list(idx) == list(range(offs[level], offs[level] + loads[level]))
    for b in np.nonzero(lev != 0xFF)[0]:
        assert np.all(pos[b] % bw == 0)
        if lev[b] < L - 1:
            ps = int(par[b])
            pb = ps // 8
            assert lev[pb] == lev[b] + 1
            assert chi[ps] == b
            d = np.array([(ps >> 2) & 1, (ps >> 1) & 1, ps & 1])
            assert np.array_equal(pos[b], 2 * pos[pb] + bw * d)
        else:
            assert par[b] == NF
    for sidx in np.nonzero(chi != NF)[0]:
        assert par[int(chi[sidx])] == sidx


@pytest.mark.parametrize("scenario", SCENARIOS)
def test_apron_geometric(scenario):
    """Every border apron entry is the cell of the finest block at the same or a coarser level
    containing the neighbour position; out-of-domain entries are own cells (dense blocks: the
    clamped interior cell; blocks placed by refinement/moves: the reference's SPREAD form)."""
    s = make_dcgrid(scenario)
    inf = s.info()
    t = s.export_tables()
    bw, L, S = inf.block_width, inf.levels, inf.sparse_levels
    A = bw + 2
    lev, pos, ci = t["block_levels"], t["block_positions"], t["cell_indices"]
    vols = [owners(t, inf, l) for l in range(L)]
    bv = bw ** 3
    g = np.array(inf.size[:3])
    mismatch = 0
    checked = 0
    for b in np.nonzero(lev != 0xFF)[0]:
        l = int(lev[b])
        for i in range(A):
            for j in range(A):
                for k in range(A):
                    inner = 0 < i <= bw and 0 < j <= bw and 0 < k <= bw
                    q = pos[b] + np.array([i - 1, j - 1, k - 1])
                    if inner:
                        want = b * bv + local_index(i - 1, j - 1, k - 1, bw)
                    elif np.any(q < 0) or np.any(q * (1 << l) >= g):
                        c = [min(max(v, 1), bw) for v in (i, j, k)]
                        if l >= S:
                            want = b * bv + local_index(c[0] - 1, c[1] - 1, c[2] - 1, bw)
                        else:
                            want = b * bv + local_index(c[0] & (bw - 1), c[1] & (bw - 1), c[2] & (bw - 1), bw)
                    else:
                        nb = int(vols[l][tuple(q)])
                        sh = int(lev[nb]) - l
                        r = (q >> sh) - pos[nb]
                        want = nb * bv + local_index(*r, bw)
                    checked += 1
                    mismatch += int(ci[b, (i * A + j) * A + k] != want)
    assert checked > 0
    assert mismatch == 0


def test_hash_map_equals_live_blocks():
    s = make_dcgrid(("cfg1_64", 4))
    inf = s.info()
    t = s.export_tables()
    from dcgrid_amd._abi import hash_map
    hm = hash_map(t, inf)
    lib = oracle.library()
    live = set()
    for b in np.nonzero(t["block_levels"] != 0xFF)[0]:
        x, y, z = (int(v) for v in t["block_positions"][b])
        live.add((int(t["block_levels"][b]), lib.orc_grid_hash(x, y, z, inf.block_width), int(b)))
    sparse_live = {e for e in live if e[0] < inf.sparse_levels}
    assert {e for e in hm if e[0] < inf.sparse_levels} == sparse_live
    # after a refill (moves happened: move_limit > 0 somewhere) every allocated block is mapped
    if any(inf.levels_list("move_limit")):
        assert hm == live


def test_refine_picks_largest_indices():
    """refineSubblocks keeps the `lim` largest subblock indices among candidates
    (fluid_simulation_dcgrid.cu:464-466): with the geometric score every unrefined subblock of
    the first adaptation scores > 1e-4, so the children of the first sparse level are the
    highest-index subblocks of the level above."""
    p = scenes.reference_params(64)
    s = oracle.OracleSim("dcgrid", (64, 64, 64), p, 1200, precision=64)
    inf = s.info()
    t = s.export_tables()
    # level-0 children exist only below the largest-index subblocks of level 1
    chi = t["child_indices"]
    o1, m1 = inf.level_offsets[1], inf.max_blocks_level[1]
    sub = chi[8 * o1: 8 * (o1 + m1)]
    refined = np.nonzero(sub != NF)[0]
    assert len(refined) == inf.block_loads[0]


# ----------------------------------------------------------------------------- closed forms
def _uniform(n=16, prec=64, **kw):
    p = scenes.reference_params(n, kind="uniform")
    p.dx, p.rdx, p.dt = 1.0, 1.0, 0.1
    p.emission_radius = 0.0
    for k, v in kw.items():
        setattr(p, k, v)
    return oracle.OracleSim("uniform", (n, n, n), p, precision=prec), p


def test_uniform_jacobi_cosine_mode():
    """One clamped (Neumann) Jacobi sweep with zero divergence multiplies the cell-centred mode
    cos(pi k (x+1/2)/N) by (4 + 2 cos(pi k/N))/6."""
    n = 16
    s, _ = _uniform(n)
    msize = s.field_size("pressure")
    for kmode in (1, 3, 7):
        x = np.arange(n)
        mode = np.cos(np.pi * kmode * (x + 0.5) / n)
        p0 = np.broadcast_to(mode[None, None, :], (n, n, n)).reshape(-1)  # x fastest
        full = np.zeros(msize, np.float32)
        full[: n ** 3] = p0
        s.import_field("pressure", full)
        s.import_field("divergence", np.zeros(msize, np.float32))
        s.run_stage(STAGE_JACOBI, 0)
        out = s.export_field_f64("t_pressure")[: n ** 3]
        lam = (4 + 2 * np.cos(np.pi * kmode / n)) / 6
        np.testing.assert_allclose(out, lam * p0.astype(np.float32).astype(np.float64), rtol=0, atol=1e-7)


def test_uniform_jacobi_symmetric():
    n = 12
    s, _ = _uniform(n)
    msize = s.field_size("pressure")
    rng = np.random.default_rng(3)

    def J(v):
        full = np.zeros(msize, np.float32)
        full[: n ** 3] = v
        s.import_field("pressure", full)
        s.import_field("divergence", np.zeros(msize, np.float32))
        s.run_stage(STAGE_JACOBI, 0)
        return s.export_field_f64("t_pressure")[: n ** 3]

    a = rng.standard_normal(n ** 3).astype(np.float32)
    b = rng.standard_normal(n ** 3).astype(np.float32)
    lhs = float(np.dot(J(a), b.astype(np.float64)))
    rhs = float(np.dot(a.astype(np.float64), J(b)))
    assert abs(lhs - rhs) <= 1e-9 * max(1.0, abs(lhs))


def test_uniform_divergence_and_gradient_linear():
    n = 16
    s, p = _uniform(n)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    a, c, e = 0.5, -0.25, 0.125
    vel = np.stack([a * x, c * y, e * z], -1).reshape(-1, 3).astype(np.float32)
    s.import_field("velocity", vel)
    s.run_stage(STAGE_DIVERGENCE)
    div = s.export_field_f64("divergence")[: n ** 3].reshape(n, n, n)
    inner = (slice(1, -1),) * 3
    np.testing.assert_allclose(div[inner], (a + c + e) * p.rdx, rtol=0, atol=1e-12)
    # closed box: sum of central divergence = boundary-adjacent values (no-slip ghosts are 0)
    u = vel.reshape(n, n, n, 3).astype(np.float64)
    flux = (u[:, :, -1, 0] - u[:, :, 0, 0]).sum() + (u[:, -1, :, 1] - u[:, 0, :, 1]).sum() + \
        (u[-1, :, :, 2] - u[0, :, :, 2]).sum()
    assert abs(div.sum() - 0.5 * p.rdx * flux) < 1e-9
    # gradient subtraction of a linear pressure: u -= (0.5 rdx)(p_r - p_l) = g in the interior
    msize = s.field_size("pressure")
    gp = np.array([0.3, -0.2, 0.1])
    pres = np.zeros(msize, np.float32)
    pres[: n ** 3] = (gp[0] * x + gp[1] * y + gp[2] * z).reshape(-1)
    s.import_field("pressure", pres)
    s.import_field("velocity", np.zeros((n ** 3, 3), np.float32))
    s.run_stage(STAGE_APPLY_PRESSURE)
    v = s.export_field_f64("velocity").reshape(n, n, n, 3)
    np.testing.assert_allclose(v[inner], np.broadcast_to(-gp * p.rdx, v[inner].shape), rtol=0, atol=1e-6)
    P3 = pres[: n ** 3].astype(np.float64).reshape(n, n, n)  # exact central differences of the input
    gx_ = -0.5 * p.rdx * (P3[1:-1, 1:-1, 2:] - P3[1:-1, 1:-1, :-2])
    np.testing.assert_allclose(v[inner][..., 0], gx_, rtol=0, atol=1e-12)


def test_uniform_restrict_mean_and_prolongate_linear():
    n = 16
    s, _ = _uniform(n)
    msize = s.field_size("divergence")
    rng = np.random.default_rng(5)
    d0 = rng.standard_normal(n ** 3).astype(np.float32)
    full = np.zeros(msize, np.float32)
    full[: n ** 3] = d0
    s.import_field("divergence", full)
    s.run_stage(STAGE_RESTRICT)
    d = s.export_field_f64("divergence")
    d1 = d[n ** 3: n ** 3 + (n // 2) ** 3]
    want = d0.astype(np.float64).reshape(n // 2, 2, n // 2, 2, n // 2, 2).mean(axis=(1, 3, 5)).reshape(-1)
    np.testing.assert_allclose(d1, want, rtol=0, atol=1e-6)
    # prolongation: 27/9/3/1 weights reproduce a linear coarse field at the fine cell centres
    m = n // 2
    zc, yc, xc = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    f = lambda X, Y, Z: 0.5 * X - 0.75 * Y + 0.25 * Z  # noqa: E731
    coarse = f(2 * xc + 1.0, 2 * yc + 1.0, 2 * zc + 1.0)  # coarse centres in fine units (x2 + 1)
    pres = np.zeros(msize, np.float32)
    pres[n ** 3: n ** 3 + m ** 3] = coarse.reshape(-1)
    s.import_field("pressure", pres)
    s.run_stage(STAGE_PROLONGATE, 0)
    fine = s.export_field_f64("pressure")[: n ** 3].reshape(n, n, n)
    zf, yf, xf = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    want = f(xf + 0.5, yf + 0.5, zf + 0.5)
    inner = (slice(1, -1),) * 3
    np.testing.assert_allclose(fine[inner], want[inner], rtol=0, atol=1e-5)


def test_uniform_advect_identity_and_integer_shift():
    n = 16
    s, p = _uniform(n)
    rng = np.random.default_rng(7)
    q = rng.random(n ** 3).astype(np.float32)
    s.import_field("density", q)
    s.import_field("velocity", np.zeros((n ** 3, 3), np.float32))
    s.advect_density()
    assert np.array_equal(s.export_field("density"), q)  # foot point = cell centre, weight 1
    # velocity of exactly one cell per step in +x: q_new(x) = q(x-1) away from the inflow face
    shift = np.zeros((n ** 3, 3), np.float32)
    shift[:, 0] = 1.0 / (p.dt * p.rdx)
    s.import_field("density", q)
    s.import_field("velocity", shift)
    s.advect_density()
    out = s.export_field_f64("density").reshape(n, n, n)
    qq = q.reshape(n, n, n).astype(np.float64)
    np.testing.assert_allclose(out[:, :, 1:], qq[:, :, :-1], rtol=0, atol=1e-6)


def test_advect_linear_field_exact():
    """Trilinear sampling reproduces a linear field at any interior foot point."""
    n = 16
    s, p = _uniform(n)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    f = lambda X, Y, Z: 0.1 * X + 0.2 * Y - 0.05 * Z + 1.0  # noqa: E731
    s.import_field("density", f(x + 0.5, y + 0.5, z + 0.5).reshape(-1).astype(np.float32))
    rng = np.random.default_rng(11)
    vel = rng.uniform(-3, 3, (n ** 3, 3)).astype(np.float32)
    s.import_field("velocity", vel)
    s.advect_density()
    out = s.export_field_f64("density").reshape(n, n, n)
    v = vel.reshape(n, n, n, 3).astype(np.float64) * p.dt * p.rdx
    want = f(x + 0.5 - v[..., 0], y + 0.5 - v[..., 1], z + 0.5 - v[..., 2])
    inner = (slice(2, -2),) * 3
    np.testing.assert_allclose(out[inner], want[inner], rtol=0, atol=1e-6)


def test_buoyancy_and_source_closed_form():
    n = 16
    w = scenes.config(0, scale=4)
    s = oracle.OracleSim(w.kind, w.size, w.params, precision=64)
    p = w.params
    rng = np.random.default_rng(2)
    rho = rng.random(n ** 3).astype(np.float32)
    tem = rng.random(n ** 3).astype(np.float32)
    s.import_field("density", rho)
    s.import_field("temperature", tem)
    s.import_field("velocity", np.zeros((n ** 3, 3), np.float32))
    s.advect_velocity()
    v = s.export_field_f64("velocity")
    want = p.dt * (p.buoy_beta * (tem.astype(np.float64) - p.t_ambient) - p.buoy_alpha * rho.astype(np.float64))
    np.testing.assert_allclose(v[:, 1], want, rtol=1e-6, atol=1e-9)
    assert np.all(v[:, 0] == 0) and np.all(v[:, 2] == 0)
    # sphere source overrides density/temperature inside the sphere after advection
    s.advect_density()
    d = s.export_field_f64("density").reshape(n, n, n)
    t = s.export_field_f64("temperature").reshape(n, n, n)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    c = np.array(p.src_center[:])
    inside = ((x + 0.5 - c[0]) ** 2 + (y + 0.5 - c[1]) ** 2 + (z + 0.5 - c[2]) ** 2) < p.src_radius ** 2
    assert inside.sum() > 0
    assert np.all(d[inside] == p.src_density)
    assert np.all(np.abs(t[inside] - p.src_temperature) <= p.src_noise * (1 + 1e-6))


def test_value_noise_lattice_and_range():
    lib = oracle.library()
    for (x, y, z) in [(0, 0, 0), (3, -2, 5), (10, 11, -7)]:
        v = lib.orc_value_noise(float(x), float(y), float(z), 99)
        assert v == np.float32(2.0) * np.float32(lib.orc_hash01(99, x, y, z, 0)) - np.float32(1.0)
    rng = np.random.default_rng(0)
    vals = [lib.orc_value_noise(*map(float, rng.uniform(-50, 50, 3)), 7) for _ in range(2000)]
    assert -1.0 <= min(vals) and max(vals) <= 1.0 and np.std(vals) > 0.1


# ----------------------------------------------------------------------------- DCGrid closed forms
def _dense_dcgrid(n=32, bw=4, prec=64):
    """All levels dense (budget = sum of full levels), no solids, closed box."""
    p = scenes.reference_params(n)
    p.block_width = bw
    p.dx, p.rdx, p.dt = 1.0, 1.0, 0.1
    p.emission_radius = 0.0
    inf = oracle.plan((n, n, n), 10 ** 7, p)
    full = sum(inf.levels_list("full_blocks_level"))
    s = oracle.OracleSim("dcgrid", (n, n, n), p, full, precision=prec)
    return s, p


def _cell_positions(s):
    """Level-local positions and levels of every cell of the field arrays."""
    inf = s.info()
    t = s.export_tables()
    bw = inf.block_width
    bv = bw ** 3
    lib = oracle.library()
    locs = np.zeros((bv, 3), np.int64)
    for x in range(bw):
        for y in range(bw):
            for z in range(bw):
                locs[lib.orc_local_index(x, y, z, bw)] = (x, y, z)
    pos = np.repeat(t["block_positions"].astype(np.int64), bv, axis=0) + np.tile(locs, (inf.max_num_blocks, 1))
    lev = np.repeat(t["block_levels"], bv)
    return pos, lev


def test_dcgrid_dense_jacobi_cosine_mode():
    s, _ = _dense_dcgrid(32, 4)
    pos, lev = _cell_positions(s)
    n0 = 32
    kmode = 3
    p = np.zeros(s.field_size("pressure"), np.float32)
    m0 = lev == 0
    p[m0] = np.cos(np.pi * kmode * (pos[m0, 1] + 0.5) / n0)  # mode along y on level 0
    s.import_field("pressure", p)
    s.import_field("divergence", np.zeros_like(p))
    s.run_stage(STAGE_JACOBI, 0)
    out = s.export_field_f64("t_pressure")
    lam = (4 + 2 * np.cos(np.pi * kmode / n0)) / 6
    np.testing.assert_allclose(out[m0], lam * p[m0].astype(np.float64), rtol=0, atol=1e-7)


def test_dcgrid_dense_restrict_and_prolongate():
    s, _ = _dense_dcgrid(32, 8)
    pos, lev = _cell_positions(s)
    rng = np.random.default_rng(4)
    d = rng.standard_normal(s.field_size("density")).astype(np.float32)
    s.import_field("density", d)
    s.run_stage(STAGE_RESTRICT, "density")
    r = s.export_field_f64("density")
    # level-1 cells are the means of their 8 level-0 children
    m0, m1 = lev == 0, lev == 1
    fine = {tuple(v): d[i] for i, v in zip(np.nonzero(m0)[0], pos[m0])}
    for i in np.nonzero(m1)[0][:200]:
        x, y, z = pos[i]
        kids = [fine[(2 * x + a, 2 * y + b, 2 * z + c)] for a in (0, 1) for b in (0, 1) for c in (0, 1)]
        assert abs(r[i] - np.mean(np.array(kids, np.float64))) < 1e-6
    # prolongation of a linear level-1 field to level 0 is exact in the interior
    f = lambda X, Y, Z: 0.3 * X - 0.1 * Y + 0.2 * Z  # noqa: E731
    pres = np.zeros(s.field_size("pressure"), np.float32)
    pres[m1] = f(2 * pos[m1, 0] + 1.0, 2 * pos[m1, 1] + 1.0, 2 * pos[m1, 2] + 1.0)
    s.import_field("pressure", pres)
    s.run_stage(STAGE_PROLONGATE, 0)
    out = s.export_field_f64("pressure")
    want = f(pos[:, 0] + 0.5, pos[:, 1] + 0.5, pos[:, 2] + 0.5)
    inner = m0 & np.all((pos >= 1) & (pos <= 30), axis=1)
    np.testing.assert_allclose(out[inner], want[inner], rtol=0, atol=1e-5)


def test_dcgrid_advect_identity():
    s = make_dcgrid(("cfg1_64", 3))
    p = s.params.copy()
    p.source_mode = 0  # no sphere override
    s.set_params(p)
    rng = np.random.default_rng(8)
    q = rng.random(s.field_size("density")).astype(np.float32)
    s.import_field("density", q)
    s.import_field("velocity", np.zeros((s.field_size("density"), 3), np.float32))
    s.advect_density()
    out = s.export_field("density")
    t = s.export_tables()
    inf = s.info()
    bv = inf.block_width ** 3
    leaf = np.repeat(t["child_indices"] == NF, bv // 8) & np.repeat(t["block_levels"] != 0xFF, bv)
    assert np.array_equal(out[leaf], q[leaf])


# ----------------------------------------------------------------------------- precision / determinism
def test_fp32_and_fp64_oracles_agree():
    w = scenes.config(1, scale=4)
    a = oracle.OracleSim(w.kind, w.size, w.params, w.max_num_blocks, precision=32)
    b = oracle.OracleSim(w.kind, w.size, w.params, w.max_num_blocks, precision=64)
    for s in (a, b):
        s.seed_scene()
        s.step(3)
    assert all(np.array_equal(x, y) for x, y in zip(a.export_tables().values(), b.export_tables().values()))
    for f in ("density", "velocity", "pressure"):
        x, y = a.export_field_f64(f), b.export_field_f64(f)
        assert np.linalg.norm(x - y) <= 1e-5 * np.linalg.norm(y)


def test_oracle_deterministic():
    r = []
    for _ in range(2):
        s = make_dcgrid(("cfg1_64", 3), precision=32)
        r.append((s.export_tables(), s.export_field("velocity")))
    for k in r[0][0]:
        assert np.array_equal(r[0][0][k], r[1][0][k])
    assert np.array_equal(r[0][1], r[1][1])


def test_create_errors():
    p = scenes.reference_params(64)
    with pytest.raises(DCGError) as e:
        oracle.OracleSim("dcgrid", (64, 64, 32), p, 1000)  # size disagrees with params
    assert e.value.code == DCG_E_ARG
    with pytest.raises(DCGError) as e:
        oracle.OracleSim("dcgrid", (64, 64, 64), p, 0)
    assert e.value.code == DCG_E_BUDGET
