"""Pins of the CPU oracle's DCGrid-only functions on adapted grids (several levels, sparse levels,
level boundaries). Each check is fixed by something outside the oracle's code:

* closed forms of linear fields (divergence, gradient, curl, trilinear advection), evaluated at
  neighbour cells found by brute-force painting of the blocks (tests/dcgrid_geometry.py) rather
  than through the oracle's apron table;
* hand-computed values of the geometric score (dcgrid_adaptation.cu:26-34);
* the reference's sequential move selection (fluid_simulation_dcgrid.cu:348-437) restated in
  numpy over the scores the oracle starts from;
* a solved pressure system, whose debugStats residual (dcgrid_structure.cu:224-251) must vanish.

CPU only.
"""
import numpy as np
import pytest

import oracle
from dcgrid_amd import scenes
from dcgrid_amd._abi import (SCORE_FEATURES, SOURCE_NONE, STAGE_APPLY_PRESSURE, STAGE_DIVERGENCE,
                             STAGE_VORTICITY)
from dcgrid_geometry import Geometry

FLT_MAX = np.float32(np.finfo(np.float32).max)


def adapted(name, steps, precision=64):
    """Oracle simulation after `steps` steps: the reference scene (BW 4, geometric scores,
    automatic levels) or configuration 1 at 64^3 (BW 8, 3 levels, feature scores)."""
    if name == "ref64_bw4":
        p = scenes.reference_params(64)
        s = oracle.OracleSim("dcgrid", (64, 64, 64), p, 1500, precision=precision)
    else:
        w = scenes.config(1, scale=4)
        s = oracle.OracleSim(w.kind, w.size, w.params, w.max_num_blocks, precision=precision)
        s.seed_scene()
    s.step(steps)
    return s


def no_sources(s):
    p = s.params.copy()
    p.source_mode = SOURCE_NONE
    p.emission_radius = 0.0
    p.velocity_emission_rate = 0.0
    p.density_emission_rate = 0.0
    s.set_params(p)
    return p


ADAPTED = [("ref64_bw4", 2), ("cfg1_64", 4)]
M = np.array([[0.5, 0.2, -0.1], [0.3, -0.25, 0.05], [-0.2, 0.15, 0.125]])
B0 = np.array([0.1, -0.3, 0.2])


def linear_velocity(G, m=M, b0=B0):
    u = np.zeros((G.B * G.bv, 3))
    u[G.alloc] = G.centre[G.alloc] @ m.T + b0
    return u


def interior(G, cells):
    nb, inside, same = G.face_owners(cells)
    return cells[inside], nb[inside], same[inside]


# ----------------------------------------------------------------------------- stencils
@pytest.mark.parametrize("scenario", ADAPTED)
def test_divergence_linear_every_level(scenario):
    """k_dcgrid_calc_divergence (dcgrid_fluid.cu:174-230) on u = M x + b: every interior leaf
    cell holds (.5 rdx/scale) sum_axis (u_a(R) - u_a(L)) at the finest block's cells that contain
    its neighbour positions, which is rdx trace(M) where those are on the cell's own level;
    non-leaf cells are zero."""
    s = adapted(*scenario)
    no_sources(s)
    G = Geometry(s)
    rdx = float(s.params.rdx)
    u = linear_velocity(G)
    s.import_field_f64("velocity", u)
    s.import_field_f64("fluidity", np.ones(G.B * G.bv))
    s.run_stage(STAGE_DIVERGENCE)
    div = s.export_field_f64("divergence")
    cells, nb, same = interior(G, np.nonzero(G.leaf)[0])
    want = 0.5 * rdx / G.scale[cells] * sum(u[nb[:, 2 * a + 1], a] - u[nb[:, 2 * a], a] for a in range(3))
    np.testing.assert_allclose(div[cells], want, rtol=0, atol=1e-12)
    np.testing.assert_allclose(div[cells[same]], rdx * np.trace(M), rtol=0, atol=1e-12)
    levels = set(G.cell_lev[cells[same]])
    assert len(levels) >= 2 and (~same).sum() > 0  # several levels and level boundaries exercised
    assert np.all(div[G.alloc & ~G.leaf] == 0.0)


@pytest.mark.parametrize("scenario", ADAPTED)
def test_gradient_linear_every_level(scenario):
    """k_dcgrid_apply_pressure (dcgrid_fluid.cu:232-259) with p = g.x + c and u = 0: every
    interior leaf cell gets -(.5 rdx/scale)(p(R) - p(L)) per axis, i.e. -rdx g where the
    neighbours are on the cell's own level; non-leaf cells keep their velocity."""
    s = adapted(*scenario)
    G = Geometry(s)
    rdx = float(s.params.rdx)
    g = np.array([0.3, -0.2, 0.45])
    pres = np.zeros(G.B * G.bv)
    pres[G.alloc] = G.centre[G.alloc] @ g + 1.5
    s.import_field_f64("pressure", pres)
    s.import_field_f64("fluidity", np.ones(G.B * G.bv))
    s.import_field_f64("velocity", np.zeros((G.B * G.bv, 3)))
    s.run_stage(STAGE_APPLY_PRESSURE)
    v = s.export_field_f64("velocity")
    cells, nb, same = interior(G, np.nonzero(G.leaf)[0])
    want = np.stack([-0.5 * rdx / G.scale[cells] * (pres[nb[:, 2 * a + 1]] - pres[nb[:, 2 * a]]) for a in range(3)], 1)
    np.testing.assert_allclose(v[cells], want, rtol=0, atol=1e-12)
    np.testing.assert_allclose(v[cells[same]], np.broadcast_to(-rdx * g, (same.sum(), 3)), rtol=0, atol=1e-12)
    assert (~same).sum() > 0
    assert np.all(v[G.alloc & ~G.leaf] == 0.0)


@pytest.mark.parametrize("scenario", ADAPTED)
def test_vorticity_linear_every_level(scenario):
    """k_dcgrid_calc_vorticity (dcgrid_fluid.cu:146-172) on u = M x + b: central curl over the
    neighbour cells, rdx (M21 - M12, M02 - M20, M10 - M01) where they are on the own level;
    computed for leaf and non-leaf cells alike."""
    s = adapted(*scenario)
    G = Geometry(s)
    rdx = float(s.params.rdx)
    u = linear_velocity(G)
    s.import_field_f64("velocity", u)
    s.import_field_f64("fluidity", np.ones(G.B * G.bv))
    s.run_stage(STAGE_VORTICITY)
    w = s.export_field_f64("vorticity")
    cells, nb, same = interior(G, np.nonzero(G.alloc)[0])
    L_, R_, D_, U_, Bk, F_ = (nb[:, k] for k in range(6))
    a = 0.5 * rdx / G.scale[cells]
    want = np.stack([a * ((u[U_, 2] - u[D_, 2]) - (u[F_, 1] - u[Bk, 1])),
                     a * ((u[F_, 0] - u[Bk, 0]) - (u[R_, 2] - u[L_, 2])),
                     a * ((u[R_, 1] - u[L_, 1]) - (u[U_, 0] - u[D_, 0]))], 1)
    np.testing.assert_allclose(w[cells], want, rtol=0, atol=1e-12)
    curl = rdx * np.array([M[2, 1] - M[1, 2], M[0, 2] - M[2, 0], M[1, 0] - M[0, 1]])
    np.testing.assert_allclose(w[cells[same]], np.broadcast_to(curl, (same.sum(), 3)), rtol=0, atol=1e-12)
    assert np.any(~G.leaf[cells[same]])  # non-leaf cells covered


# ----------------------------------------------------------------------------- cross-level advection
def feet_and_corners(G, cells, u, dt, rdx):
    """Foot points of the cells (finest units) and, per cell, whether the 2x2x2 sample stencil
    of the finest block containing the foot consists of in-domain cells of that block's level
    (trilinear weights then reproduce a linear field exactly); also that block's level."""
    foot = G.centre[cells] - u[cells] * (dt * rdx)
    ipos = np.clip(np.floor(foot).astype(np.int64), 0, G.size - 1)
    blk = G.owners[0][ipos[:, 0], ipos[:, 1], ipos[:, 2]]
    lb = G.t["block_levels"][blk].astype(np.int64)
    x = foot / (1 << lb)[:, None] - 0.5
    x0 = np.floor(x).astype(np.int64)
    ok = np.ones(len(cells), bool)
    for k in range(8):
        corner = x0 + np.array([(k >> 2) & 1, (k >> 1) & 1, k & 1])
        good = np.zeros(len(cells), bool)
        for level in range(G.L):
            m = lb == level
            if m.any():
                _, o, ol = G.owner_cells(corner[m], level)
                good[m] = o & (ol == level)
        ok &= good
    return foot, ok, lb


@pytest.mark.parametrize("scenario", ADAPTED)
def test_advect_density_linear_across_levels(scenario):
    """INIT_SAMPLE + k_dcgrid_advect_density (dcgrid_fluid.cu:7-72,132-144): a linear density
    advected by random velocities is reproduced exactly, f(foot), at every leaf cell whose foot's
    sample stencil lies on one level — including feet that land in finer or coarser blocks than
    the cell's own (getBlockIndexDeep)."""
    s = adapted(*scenario)
    p = no_sources(s)
    G = Geometry(s)
    N = G.B * G.bv
    f = lambda X: 0.02 * X[:, 0] - 0.015 * X[:, 1] + 0.01 * X[:, 2] + 0.5  # noqa: E731
    q = np.zeros(N)
    q[G.alloc] = f(G.centre[G.alloc])
    rng = np.random.default_rng(12)
    u = rng.uniform(-1.0, 1.0, (N, 3)) * (2.5 / (float(p.dt) * float(p.rdx)))  # feet up to 2.5 cells away
    s.import_field_f64("density", q)
    s.import_field_f64("velocity", u)
    s.import_field_f64("fluidity", np.ones(N))
    s.advect_density()
    out = s.export_field_f64("density")
    cells = np.nonzero(G.leaf)[0]
    foot, ok, lb = feet_and_corners(G, cells, u, float(p.dt), float(p.rdx))
    np.testing.assert_allclose(out[cells[ok]], f(foot[ok]), rtol=0, atol=1e-9)
    cross = lb[ok] != G.cell_lev[cells[ok]]
    assert ok.sum() > 0.5 * len(cells) and cross.sum() > 20  # feet in blocks of other levels exercised


@pytest.mark.parametrize("scenario", ADAPTED)
def test_advect_velocity_linear_across_levels(scenario):
    """k_dcgrid_advect_velocity (dcgrid_fluid.cu:112-130): a linear velocity field advects
    itself to u(foot) = M foot + b at every leaf cell whose foot's stencil lies on one level."""
    s = adapted(*scenario)
    p = no_sources(s)
    G = Geometry(s)
    N = G.B * G.bv
    sc = 3.0 / (float(p.dt) * float(p.rdx) * 64.0)  # |u| dt rdx up to about 3 cells
    m = M * sc
    b0 = B0 * sc
    u = linear_velocity(G, m, b0)
    s.import_field_f64("velocity", u)
    s.import_field_f64("fluidity", np.ones(N))
    s.advect_velocity()
    out = s.export_field_f64("velocity")
    cells = np.nonzero(G.leaf)[0]
    foot, ok, lb = feet_and_corners(G, cells, u, float(p.dt), float(p.rdx))
    np.testing.assert_allclose(out[cells[ok]], foot[ok] @ m.T + b0, rtol=0, atol=1e-9)
    assert ok.sum() > 0.5 * len(cells) and (lb[ok] != G.cell_lev[cells[ok]]).sum() > 0


# ----------------------------------------------------------------------------- scores
def geometric_scores_np(G, gx, gy, gz):
    """k_dcgrid_calc_subblock_scores / k_dcgrid_accumulate_subblock_scores
    (dcgrid_adaptation.cu:10-63) in float32, evaluated as the reference writes them; the
    subblock centre generalises (pos + 2 d + 1) of 2^3-cell subblocks to (pos + (BW/2) d + BW/4)."""
    t, inf = G.t, G.inf
    B, bw, L = G.B, G.bw, G.L
    lev = t["block_levels"].astype(np.int64)
    loads = inf.levels_list("block_loads")
    full = inf.levels_list("full_blocks_level")
    blocked_level = np.array([(l == 0 and loads[0] == full[0]) or (l > 0 and loads[l - 1] == full[l - 1])
                              for l in range(L)] + [True])
    s = np.arange(8 * B)
    b = s // 8
    blocked = (t["child_indices"] != np.uint64(0xFFFFFFFFFFFFFFFF)) | (lev[b] == 0xFF) | blocked_level[np.minimum(lev[b], L)]
    f32 = np.float32
    sc = (np.int64(1) << np.where(lev[b] == 0xFF, 0, lev[b])).astype(f32)
    hb, qb = f32(bw // 2), f32(bw // 4)
    pos = t["block_positions"][b].astype(f32)
    bits = np.stack([(s >> 2) & 1, (s >> 1) & 1, s & 1], 1).astype(f32)
    q = sc[:, None] * ((pos + hb * bits) + qb)
    c = np.array([f32(.5) * f32(gx), f32(.45) * f32(gy), f32(.5) * f32(gz)], f32)
    dxyz = q - c
    d = np.sqrt((dxyz[:, 0] * dxyz[:, 0] + dxyz[:, 1] * dxyz[:, 1]) + dxyz[:, 2] * dxyz[:, 2])
    with np.errstate(divide="ignore"):
        sub = np.where(d < f32(.2) * f32(gx), f32(0), f32(gx) / d).astype(f32)
    sub = np.where(blocked, -FLT_MAX, sub).astype(f32)
    sub = np.append(sub, -FLT_MAX).astype(f32)
    blk = np.full(B, -FLT_MAX, f32)
    okb = (lev != 0xFF) & ~blocked_level[np.minimum(lev, L)]
    s8 = sub[:8 * B].reshape(B, 8)
    allpos = np.all(s8 > 0, axis=1)
    acc = sub[0:8 * B:8].copy()
    with np.errstate(over="ignore"):  # -FLT_MAX terms overflow only in blocks that are masked out
        for k in range(1, 9):
            acc = (acc + sub[k:8 * B + k:8]).astype(f32)  # s[k] of block b is sub[8b + k]; k = 8: next block
    blk = np.where(okb & allpos, (f32(.125) * acc).astype(f32), blk)
    return sub, blk


def test_geometric_score_hand_values():
    """Hand-computed subblock scores of the reference scene: centre p = 2^l (pos + 2 d + 1),
    distance to (.5 gx, .45 gy, .5 gz), score gx/d beyond .2 gx, else 0."""
    s = adapted("ref64_bw4", 0)
    G = Geometry(s)
    sub, _ = s.scores()
    t = G.t
    lev = t["block_levels"]
    checked = zeros = 0
    import math
    for sidx in range(0, 8 * G.B, 37):
        b = sidx // 8
        if lev[b] == 0xFF or t["child_indices"][sidx] != np.uint64(0xFFFFFFFFFFFFFFFF) or sub[sidx] == -FLT_MAX:
            continue
        l = int(lev[b])
        d3 = ((sidx >> 2) & 1, (sidx >> 1) & 1, sidx & 1)
        p = [(1 << l) * (int(t["block_positions"][b][a]) + 2 * d3[a] + 1) for a in range(3)]
        d = math.sqrt((p[0] - 32.0) ** 2 + (p[1] - 28.8) ** 2 + (p[2] - 32.0) ** 2)
        want = 0.0 if d < 12.8 else 64.0 / d
        assert abs(float(sub[sidx]) - want) <= 2e-6 * max(want, 1.0), (sidx, sub[sidx], want)
        checked += 1
        zeros += want == 0.0
    assert checked > 20 and zeros > 0


@pytest.mark.parametrize("scenario", [("ref64_bw4", 0), ("ref64_bw4", 3)])
def test_geometric_scores_float32_order(scenario):
    """The oracle's subblock and block scores equal the float32 evaluation of the reference
    expressions bit for bit (including the -FLT_MAX cases and the block score's ninth term)."""
    s = adapted(*scenario)
    G = Geometry(s)
    sub, blk = s.scores()
    want_sub, want_blk = geometric_scores_np(G, 64, 64, 64)
    assert np.array_equal(sub, want_sub)
    assert np.array_equal(blk, want_blk)
    assert np.sum(blk > 0) > 10


def test_feature_score_linear_fields():
    """Feature score (extension, configuration 1): max over a subblock's cells of
    max(|grad rho| / grad_threshold, |curl u| / vort_threshold) with central differences; for
    rho = g.x and u = M x it is max(rdx |g| / gt, rdx |curl M| / wt) on subblocks whose cells
    have their neighbours on their own level."""
    s = adapted("cfg1_64", 3)
    assert s.params.score_mode == SCORE_FEATURES
    G = Geometry(s)
    rdx = float(s.params.rdx)
    for g, m in [(np.array([0.02, -0.01, 0.015]), M * 0.01), (np.array([0.2, 0.1, -0.3]), M * 1e-4)]:
        rho = np.zeros(G.B * G.bv)
        rho[G.alloc] = G.centre[G.alloc] @ g
        s.import_field_f64("density", rho)
        s.import_field_f64("velocity", linear_velocity(G, m, B0))
        s.import_field_f64("fluidity", np.ones(G.B * G.bv))
        sub, _ = s.scores()
        cells, nb, same = interior(G, np.nonzero(G.leaf)[0])
        per_sub = G.bv // 8
        good = np.zeros(8 * G.B, np.int64)
        np.add.at(good, cells[same] // per_sub, 1)
        full = np.nonzero(good == per_sub)[0]
        assert len(full) > 10
        curl = np.array([m[2, 1] - m[1, 2], m[0, 2] - m[2, 0], m[1, 0] - m[0, 1]])
        want = max(rdx * np.linalg.norm(g) / float(s.params.grad_threshold),
                   rdx * np.linalg.norm(curl) / float(s.params.vort_threshold))
        sel = full[sub[full] != -FLT_MAX]
        assert len(sel) > 10
        np.testing.assert_allclose(sub[sel], want, rtol=2e-6)


# ----------------------------------------------------------------------------- move selection
def select_moves_np(sub, blk, inf):
    """moveBlocks' selection (fluid_simulation_dcgrid.cu:348-437) over the given scores, with
    the total orders that replace the reference's comparator: candidates (non-negative first,
    score ascending, index ascending), destinations (score descending, index ascending). Returns
    {block: destination subblock}, the new move limits and (limit, matches) per level."""
    bs = blk.copy()
    L = int(inf.levels)
    maxL = inf.levels_list("max_blocks_level")
    full = inf.levels_list("full_blocks_level")
    offs = inf.levels_list("level_offsets")
    loads = inf.levels_list("block_loads")
    ml = list(inf.levels_list("move_limit"))
    moves, counts = {}, []
    for l in range(L - 1):
        d0, d1 = maxL[l], 8 * maxL[l + 1]
        lim = min(d0, d1, loads[l], full[l] - loads[l])
        if ml[l] > 0:
            lim = min(lim, ml[l])
        if lim == 0:
            counts.append((0, 0))
            continue
        mc = offs[l] + np.arange(d0)
        sc = bs[mc]
        mc = mc[np.lexsort((mc, np.where(sc < 0, np.float32(0), sc), sc < 0))]
        dc = 8 * offs[l + 1] + np.arange(d1)
        dc = dc[np.lexsort((dc, -sub[dc].astype(np.float64)))]
        m = 0
        while m < lim and bs[mc[m]] >= 0 and sub[dc[m]] >= 0 and bs[mc[m]] < sub[dc[m]]:
            m += 1
            if m < d1:
                bs[dc[m] // 8] = -FLT_MAX  # the reference marks the block of the next destination
        for i in range(m):
            moves[int(mc[i])] = int(dc[i])
        ml[l] = int(np.float32(m) * np.float32(1.2))
        counts.append((lim, m))
    return moves, ml, counts


def hot_spot_state(steps, seed):
    """Configuration 1 at 64^3 after `steps` steps, move limits cleared and a noisy density blob
    at a seeded place: many blocks want to move at once, so the level limits bind."""
    s = adapted("cfg1_64", steps)
    inf = s.info()
    s.import_tables(s.export_tables(), inf.levels_list("block_loads"), [0] * inf.levels)
    G = Geometry(s)
    rng = np.random.default_rng(seed)
    cx, r = rng.uniform(8, 56, 3), rng.uniform(6, 20)
    d = np.linalg.norm(G.centre - cx, axis=1)
    s.import_field_f64("density", np.where(d < r, rng.random(G.B * G.bv) * rng.uniform(0.1, 3), 0.0))
    return s


MOVE_STATES = [("ref64_bw4", 0), ("cfg1_64", 2), ("cfg1_64", 5), ("hot", 1, 1), ("hot", 1, 4), ("hot", 2, 6),
               ("hot", 6, 9), ("hot", 6, 2)]


def move_state(key):
    return hot_spot_state(key[1], key[2]) if key[0] == "hot" else adapted(*key)


@pytest.mark.parametrize("key", MOVE_STATES)
def test_move_selection_bruteforce(key):
    s = move_state(key)
    inf = s.info()
    t0 = s.export_tables()
    sub, blk = s.scores()
    moves, ml, _ = select_moves_np(sub, blk, inf)
    s.adapt_topology()
    t1 = s.export_tables()
    inf1 = s.info()
    moved = np.nonzero((t1["block_flags"] & 1) & (t0["block_levels"] != 0xFF))[0]
    got = {int(b): int(t1["parent_indices"][b]) for b in moved}
    assert got == moves
    assert inf1.levels_list("move_limit") == ml
    # the moved blocks sit under their new parent subblock
    bw = inf.block_width
    for b, d in moves.items():
        pb = d // 8
        off = np.array([(d >> 2) & 1, (d >> 1) & 1, d & 1]) * bw
        assert np.array_equal(t1["block_positions"][b], 2 * t1["block_positions"][pb] + off)


def test_move_selection_exercised():
    """The brute-force states move blocks, on both sparse levels, some of them most of what the
    level limit min(B_l, 8 B_(l+1), load, full - load) allows."""
    total, near, levels = 0, 0, set()
    for key in MOVE_STATES:
        s = move_state(key)
        sub, blk = s.scores()
        moves, _, counts = select_moves_np(sub, blk, s.info())
        total += len(moves)
        for l, (lim, m) in enumerate(counts):
            if m:
                levels.add(l)
            near += int(lim > 0 and m >= 0.8 * lim)
    assert total > 40 and levels == {0, 1} and near > 0


# ----------------------------------------------------------------------------- residual
@pytest.mark.parametrize("scenario", [("ref64_bw4", 2), ("cfg1_64", 3)])
def test_residual_after_converged_solve(scenario):
    """debugStats (dcgrid_structure.cu:224-251): sum over blocks of scale * |div - lap(p)| at leaf
    cells. After the oracle's SOR solve of the level systems to 1e-12 it vanishes on every level
    with coarser ghosts; a level that covers the domain (pure Neumann) is solved for its
    divergence minus its mean, so each of its leaf cells keeps |mean| exactly."""
    s = adapted(*scenario)
    s.project_solve()
    r_solved = s.debug_stats()
    G = Geometry(s)
    inf = s.info()
    div = s.export_field_f64("divergence")
    loads, mx, full = (inf.levels_list(k) for k in ("block_loads", "max_blocks_level", "full_blocks_level"))
    want = 0.0
    sparse_leaves = 0
    for l in range(G.L):
        lv = G.cell_lev == l
        if loads[l] == mx[l] == full[l]:
            want += (1 << l) * abs(div[G.alloc & lv].mean()) * np.sum(G.leaf & lv)
        else:
            sparse_leaves += np.sum(G.leaf & lv)
    s.import_field_f64("pressure", np.zeros(s.field_size("pressure")))
    r_zero = s.debug_stats()
    assert r_zero > 0 and sparse_leaves > 1000
    assert abs(r_solved - want) <= 1e-6 * r_zero, (r_solved, want, r_zero)
