"""GPU parity of the differentiable rollout (uniform grid) through the C ABI: the HIP adjoint
against the FP64 oracle adjoint (the reference-logic gradient), against FP64 central finite
differences at 16 seeded coordinates (the 32^3, 8-step control check), checkpoint-interval
invariance, the torch.autograd binding, and error paths."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dcgrid_amd  # noqa: E402
from dcgrid_amd import scenes  # noqa: E402
from dcgrid_amd._abi import PROJECT_SOLVE  # noqa: E402
from parity_util import rel_l2  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def torch_cuda():
    import torch
    assert torch.cuda.is_available()
    return torch


def control_case(n):
    w = scenes.config(3, scale=128 // n)
    den, vel, tgt = scenes.control_inputs(n)
    return w, den, vel, tgt


def plume_case(n, solids):
    w = scenes.config(0, scale=64 // n)
    w.params.enable_additional_solids = 1 if solids else 0
    w.params.project_mode = dcgrid_amd.PROJECT_MULTILEVEL
    w.params.project_top_pairs, w.params.project_level_pairs = 2, 1
    return w


def gpu_gradient(w, steps, k, den=None, vel=None, seed_scene=False, gd=None, gv=None, gt=None):
    """Run the rollout on the GPU and pull back the seeds; returns (d_den, d_vel, d_tem, final
    state (den, vel, tem))."""
    torch = torch_cuda()
    sim = dcgrid_amd.Simulation("uniform", w.size, w.params)
    if seed_scene:
        sim.seed_scene()
    if den is not None:
        sim.import_field("density", den)
    if vel is not None:
        sim.import_field("velocity", vel)
    sim.tape_forward(steps, k)
    fin = (sim.export_field("density"), sim.export_field("velocity"), sim.export_field("temperature"))
    N = fin[0].size
    dev = torch.device("cuda")
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)  # noqa: E731
    g_d = t(gd(fin) if callable(gd) else gd)
    g_v = t(gv if gv is not None else np.zeros((N, 3)))
    g_t = t(gt) if gt is not None else None
    o_d = torch.empty(N, device=dev)
    o_v = torch.empty(N, 3, device=dev)
    o_t = torch.empty(N, device=dev)
    torch.cuda.synchronize()
    sim.tape_backward(g_d.data_ptr(), g_v.data_ptr(), g_t.data_ptr() if g_t is not None else None,
                      o_d.data_ptr(), o_v.data_ptr(), o_t.data_ptr())
    after = (sim.export_field("density"), sim.export_field("velocity"))
    sim.close()
    return o_d.cpu().numpy(), o_v.cpu().numpy(), o_t.cpu().numpy(), fin, after


def oracle_gradient(w, steps, den=None, vel=None, seed_scene=False, gd=None, gv=None, gt=None, precision=64):
    import oracle
    s = oracle.OracleSim("uniform", w.size, w.params, precision=precision)
    if seed_scene:
        s.seed_scene()
    if den is not None:
        s.import_field("density", den)
    if vel is not None:
        s.import_field("velocity", vel)
    s.tape_forward(steps)
    fin = (s.export_field_f64("density"), s.export_field_f64("velocity"), s.export_field_f64("temperature"))
    N = fin[0].size
    g_d = gd(fin) if callable(gd) else gd
    out = s.tape_backward(g_d, gv if gv is not None else np.zeros((N, 3)), gt)
    s.close()
    return out


# The FP32 rollout is bit-identical to the oracle's FP32 instance, so the adjoints are compared
# with it at 1e-5 (only summation order differs: gathers here, scatters there). Against the FP64
# instance the gradient also carries the FP32 forward's own rounding (the FP32 oracle shows the
# same 3.2e-4 velocity-gradient distance on the control case), hence 1e-3 there.
TOL_FP32, TOL_FP64 = 1e-5, 1e-3


def test_control32_vs_oracle_adjoint():
    """L = 0.5 |rho_8 - target|^2 on the 32^3 control variant; checkpoints every 3 steps (a
    partial last segment)."""
    w, den, vel, tgt = control_case(32)
    say("[case] control32 K=3")
    seed = lambda fin: fin[0] - tgt  # noqa: E731
    gd, gv, _, _, _ = gpu_gradient(w, 8, 3, den, vel, gd=seed)
    for prec, tol in ((32, TOL_FP32), (64, TOL_FP64)):
        od, ov, _ = oracle_gradient(w, 8, den, vel, gd=seed, precision=prec)
        say(f"fp{prec} rel l2 density", rel_l2(gd, od), "velocity", rel_l2(gv, ov))
        assert rel_l2(gd, od) <= tol
        assert rel_l2(gv, ov) <= tol


@pytest.mark.parametrize("solids", [False, True])
def test_plume_with_sources_vs_oracle_adjoint(solids):
    """Buoyant plume with the emitting sphere (temperature, buoyancy, source masking), with and
    without solids (fluidity-weighted sampling, divergence and gradient): random seeds on all
    three final fields."""
    w = plume_case(16, solids)
    say(f"[case] plume16 solids={solids}")
    rng = np.random.Generator(np.random.PCG64(5))
    N = 16 ** 3
    gd0, gv0, gt0 = rng.normal(size=N), rng.normal(size=(N, 3)), rng.normal(size=N)
    gd, gv, gt, _, _ = gpu_gradient(w, 3, 2, seed_scene=True, gd=gd0, gv=gv0, gt=gt0)
    for prec, tol in ((32, TOL_FP32), (64, TOL_FP64)):
        od, ov, ot = oracle_gradient(w, 3, seed_scene=True, gd=gd0, gv=gv0, gt=gt0, precision=prec)
        say(f"fp{prec} rel l2", rel_l2(gd, od), rel_l2(gv, ov), rel_l2(gt, ot))
        assert rel_l2(gd, od) <= tol
        assert rel_l2(gv, ov) <= tol
        assert rel_l2(gt, ot) <= tol


def test_control32_vs_fp64_finite_differences():
    """The control gradient check: GPU adjoint vs FP64 central differences (h = 1e-4) of the
    oracle's forward at 16 seeded coordinates of u0 and of rho0; relative error of each
    16-vector <= 1e-4."""
    import oracle
    w, den, vel, tgt = control_case(32)
    say("[case] control32 FD")
    gd, gv, _, _, _ = gpu_gradient(w, 8, 8, den, vel, gd=lambda fin: fin[0] - tgt)
    d64, v64 = den.astype(np.float64), vel.astype(np.float64)

    def loss(d0, v0):
        s = oracle.OracleSim("uniform", w.size, w.params, precision=64)
        s.import_field_f64("density", d0)
        s.import_field_f64("velocity", v0)
        s.step(8)
        d = s.export_field_f64("density")
        s.close()
        return 0.5 * np.sum((d - tgt) ** 2)

    rng = np.random.Generator(np.random.PCG64(17))
    cand = np.argsort(-np.abs(gv).max(axis=1))[:2000]
    h = 1e-4
    fv, av, fd_, ad = [], [], [], []
    for _ in range(16):
        i = int(rng.choice(cand))
        a = int(rng.integers(3))
        vp, vm = v64.copy(), v64.copy()
        vp[i, a] += h
        vm[i, a] -= h
        fv.append((loss(d64, vp) - loss(d64, vm)) / (2 * h))
        av.append(gv[i, a])
        dp, dm = d64.copy(), d64.copy()
        dp[i] += h
        dm[i] -= h
        fd_.append((loss(dp, v64) - loss(dm, v64)) / (2 * h))
        ad.append(gd[i])
    fv, av, fd_, ad = map(np.array, (fv, av, fd_, ad))
    say("FD rel err velocity", np.linalg.norm(av - fv) / np.linalg.norm(fv),
        "density", np.linalg.norm(ad - fd_) / np.linalg.norm(fd_))
    assert np.linalg.norm(av - fv) <= 1e-4 * np.linalg.norm(fv)
    assert np.linalg.norm(ad - fd_) <= 1e-4 * np.linalg.norm(fd_)


def test_checkpoint_interval_invariance_and_final_state():
    """Recomputation from checkpoints is exact: intervals 1, 3 and 8 give bitwise identical
    gradients, and after the backward pass the live state is the rollout's final state."""
    w, den, vel, tgt = control_case(32)
    say("[case] checkpoint invariance")
    seed = lambda fin: fin[0] - tgt  # noqa: E731
    res = [gpu_gradient(w, 8, k, den, vel, gd=seed) for k in (1, 3, 8)]
    for r in res[1:]:
        assert np.array_equal(r[0], res[0][0])
        assert np.array_equal(r[1], res[0][1])
    for r in res:
        fin, after = r[3], r[4]
        assert np.array_equal(after[0], fin[0])
        assert np.array_equal(after[1], fin[1])


def test_autograd_function_and_descent():
    """torch.autograd binding: the gradient of the loss equals the direct C-ABI adjoint, and
    three steps of gradient descent on u0 lower the loss."""
    torch = torch_cuda()
    from dcgrid_amd.autograd import rollout
    w, den, vel, tgt = control_case(32)
    say("[case] autograd")
    sim = dcgrid_amd.Simulation("uniform", w.size, w.params)
    dev = torch.device("cuda")
    d0 = torch.as_tensor(den, device=dev)
    v0 = torch.as_tensor(vel, device=dev).requires_grad_(True)
    target = torch.as_tensor(tgt, device=dev)
    dN, _ = rollout(sim, d0, v0, 8, 4)
    loss = 0.5 * torch.sum((dN - target) ** 2)
    loss.backward()
    ref = gpu_gradient(w, 8, 4, den, vel, gd=lambda fin: fin[0] - tgt)
    assert np.array_equal(v0.grad.cpu().numpy(), ref[1])
    losses = [loss.item()]
    v = v0.detach().clone()
    grad = v0.grad
    for _ in range(3):
        # steepest descent with the largest velocity change 0.01 per iteration
        v = (v - (0.01 / float(grad.abs().max())) * grad).detach().requires_grad_(True)
        dN, _ = rollout(sim, d0, v, 8, 4)
        loss = 0.5 * torch.sum((dN - target) ** 2)
        loss.backward()
        losses.append(loss.item())
        grad = v.grad
    say("losses", losses)
    assert losses[-1] < losses[0]
    sim.close()


def test_adjoint_errors():
    """The projection solve has no adjoint (DCG_E_UNSUPPORTED); backward before forward is an
    error."""
    w = scenes.config(0, scale=4)
    w.params.project_mode = PROJECT_SOLVE
    sim = dcgrid_amd.Simulation(w.kind, w.size, w.params)
    with pytest.raises(dcgrid_amd.DCGError) as e:
        sim.tape_forward(2, 1)
    assert e.value.code == -4
    sim.close()
    torch = torch_cuda()
    w3, _, _, _ = control_case(32)
    N = 32 ** 3
    sim = dcgrid_amd.Simulation("uniform", w3.size, w3.params)
    a = torch.zeros(N, device="cuda")
    b = torch.zeros(N, 3, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(dcgrid_amd.DCGError):  # no recorded rollout: nothing is read or written
        sim.tape_backward(a.data_ptr(), b.data_ptr(), None, a.data_ptr(), b.data_ptr(), None)
    sim.close()
