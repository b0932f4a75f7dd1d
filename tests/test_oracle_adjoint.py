"""Pins of the oracle's reverse-mode derivative of the uniform-grid step (FP64), the
reference-logic gradient the GPU adjoint is compared with: central finite differences of the
oracle's own forward step, the dot-product identity <J x, y> = <x, J^T y>, and the closed form
of a motionless fluid (the rollout is the identity on density)."""
import numpy as np
import pytest

from dcgrid_amd import scenes


def control_case(n=32):
    w = scenes.config(3, scale=128 // n)
    den, vel, tgt = scenes.control_inputs(n)
    return w, den.astype(np.float64), vel.astype(np.float64), tgt.astype(np.float64)


def plume_case(n=16, solids=False, local=False):
    w = scenes.config(0, scale=64 // n)
    p = w.params
    p.enable_additional_solids = 1 if solids else 0
    if local:
        p.project_local_pairs = 3
    else:
        p.project_mode = 0
        p.project_top_pairs, p.project_level_pairs = 2, 1
    return w


def rollout(oracle_lib, w, steps, den=None, vel=None, tem=None, tape=False, seed_scene=False):
    s = oracle_lib.OracleSim("uniform", w.size, w.params, precision=64)
    if seed_scene:
        s.seed_scene()
    if den is not None:
        s.import_field_f64("density", den)
    if vel is not None:
        s.import_field_f64("velocity", vel)
    if tem is not None:
        s.import_field_f64("temperature", tem)
    if tape:
        s.tape_forward(steps)
    else:
        s.step(steps)
    return s


def final(s):
    return s.export_field_f64("density"), s.export_field_f64("velocity"), s.export_field_f64("temperature")


def test_fd_16_coordinates_control_32(oracle_lib):
    """The config-3 gradient check: d L / d u0 and d L / d rho0 of L = 0.5 |rho_8 - target|^2 on
    the 32^3, 8-step variant, against FP64 central differences (h = 1e-4) at 16 seeded
    coordinates each: relative error of the 16-vector <= 1e-4."""
    w, den, vel, tgt = control_case(32)
    steps = 8
    s = rollout(oracle_lib, w, steps, den, vel, tape=True)
    dN, _, _ = final(s)
    gd, gv, _ = s.tape_backward(dN - tgt, np.zeros_like(vel))
    s.close()

    def loss(d0, v0):
        r = rollout(oracle_lib, w, steps, d0, v0)
        d = r.export_field_f64("density")
        r.close()
        return 0.5 * np.sum((d - tgt) ** 2)

    rng = np.random.Generator(np.random.PCG64(7))
    N = den.size
    h = 1e-4
    fd_v, ad_v, fd_d, ad_d = [], [], [], []
    # coordinates near the blob where the gradient is not negligible
    cand = np.argsort(-np.abs(gv).max(axis=1))[:2000]
    for _ in range(16):
        i = int(rng.choice(cand))
        a = int(rng.integers(3))
        vp, vm = vel.copy(), vel.copy()
        vp[i, a] += h
        vm[i, a] -= h
        fd_v.append((loss(den, vp) - loss(den, vm)) / (2 * h))
        ad_v.append(gv[i, a])
        dp, dm = den.copy(), den.copy()
        dp[i] += h
        dm[i] -= h
        fd_d.append((loss(dp, vel) - loss(dm, vel)) / (2 * h))
        ad_d.append(gd[i])
    fd_v, ad_v, fd_d, ad_d = map(np.array, (fd_v, ad_v, fd_d, ad_d))
    assert np.linalg.norm(ad_v - fd_v) <= 1e-4 * np.linalg.norm(fd_v)
    assert np.linalg.norm(ad_d - fd_d) <= 1e-4 * np.linalg.norm(fd_d)


@pytest.mark.parametrize("case", ["control", "plume_sources", "plume_solids", "plume_local"])
def test_dot_product_identity(oracle_lib, case):
    """<J x, y> (central difference of the FP64 forward along x) equals <x, J^T y> (the adjoint)
    for random x over (density, velocity, temperature) and random y over the same outputs."""
    if case == "control":
        w, den, vel, _ = control_case(16)
        tem, seed_scene = None, False
        steps = 4
    else:
        w = plume_case(16, solids=case == "plume_solids", local=case == "plume_local")
        den = vel = tem = None
        seed_scene = True
        steps = 3
    base = rollout(oracle_lib, w, steps, den, vel, tem, tape=True, seed_scene=seed_scene)
    # initial state of the recorded rollout
    s0 = rollout(oracle_lib, w, 0, den, vel, tem, seed_scene=seed_scene)
    d0, v0, t0 = final(s0)
    s0.close()
    rng = np.random.Generator(np.random.PCG64(11))
    N = d0.size
    yd, yv, yt = rng.normal(size=N), rng.normal(size=(N, 3)), rng.normal(size=N)
    xd, xv, xt = rng.normal(size=N), rng.normal(size=(N, 3)), rng.normal(size=N)
    gd, gv, gt = base.tape_backward(yd, yv, yt)
    base.close()
    adj = np.sum(gd * xd) + np.sum(gv * xv) + np.sum(gt * xt)

    def phi(t):
        r = rollout(oracle_lib, w, steps, d0 + t * xd, v0 + t * xv, t0 + t * xt)
        d, v, tt = final(r)
        r.close()
        return np.sum(d * yd) + np.sum(v * yv) + np.sum(tt * yt)

    # The rollout is piecewise smooth: trilinear sampling has kinks where a foot crosses a cell
    # centre plane, and a small fluid velocity keeps every foot near one. The identity holds on
    # each smooth piece, so it is checked at the first of several steps whose central
    # difference does not straddle a kink (two neighbouring steps then agree).
    errs = []
    for eps in (1e-6, 1e-7, 1e-8):
        jvp = (phi(eps) - phi(-eps)) / (2 * eps)
        errs.append(abs(jvp - adj) / max(abs(adj), 1.0))
    assert min(errs) <= 1e-6, errs


def test_motionless_fluid_is_identity_on_density(oracle_lib):
    """u0 = 0 and no source: the velocity stays 0, every foot is its own cell centre and the
    rollout is the identity on density, so d rho_N / d rho_0 = I exactly and the temperature
    adjoint passes through unchanged."""
    w, den, vel, _ = control_case(16)
    s = rollout(oracle_lib, w, 3, den, np.zeros_like(vel), tape=True)
    rng = np.random.Generator(np.random.PCG64(3))
    N = den.size
    yd, yt = rng.normal(size=N), rng.normal(size=N)
    gd, gv, gt = s.tape_backward(yd, np.zeros((N, 3)), yt)
    s.close()
    assert np.array_equal(gd, yd)
    assert np.array_equal(gt, yt)
    assert np.all(np.isfinite(gv))


def test_backward_without_tape_is_an_error(oracle_lib):
    w, den, vel, _ = control_case(16)
    s = oracle_lib.OracleSim("uniform", w.size, w.params, precision=64)
    with pytest.raises(Exception):
        s.tape_backward(np.zeros(den.size), np.zeros((den.size, 3)))
    s.close()
