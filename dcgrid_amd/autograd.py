"""torch.autograd binding of the differentiable rollout (uniform grid; DCGrid on a frozen block
topology, FP32 field storage).

`rollout(sim, density0, velocity0, steps, checkpoint_every)` returns the density and velocity
after `steps` steps of the simulation step; its backward runs the hand-written adjoint kernels
of libdcgrid_amd through the C ABI (dcg_tape_forward / dcg_tape_backward). Tensors are float32
CUDA tensors: density [N] and velocity [N, 3] in the grid's cell order (uniform: x fastest;
DCGrid: block-major, the order of dcg_export_field). Temperature is held at the simulation's
current field and receives no gradient here.
"""
import torch


def _sync():
    torch.cuda.synchronize()


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, density0, velocity0, sim, steps, checkpoint_every):
        for t in (density0, velocity0):
            if not (t.is_cuda and t.dtype == torch.float32):
                raise TypeError("rollout expects float32 CUDA tensors")
        d0 = density0.detach().contiguous()
        v0 = velocity0.detach().contiguous()
        n = d0.numel()
        if v0.numel() != 3 * n:
            raise ValueError("velocity must hold 3 floats per cell")
        _sync()
        sim.import_field_device("density", d0.data_ptr(), n)
        sim.import_field_device("velocity", v0.data_ptr(), 3 * n)
        sim.tape_forward(steps, checkpoint_every)
        den = torch.empty_like(d0)
        vel = torch.empty_like(v0)
        sim.export_field_device("density", den.data_ptr(), n)
        sim.export_field_device("velocity", vel.data_ptr(), 3 * n)
        ctx.sim = sim
        ctx.shapes = (d0.shape, v0.shape)
        return den, vel

    @staticmethod
    def backward(ctx, g_density, g_velocity):
        sim = ctx.sim
        ds, vs = ctx.shapes
        dev = g_density.device if g_density is not None else g_velocity.device
        gd = g_density.contiguous() if g_density is not None else torch.zeros(ds, dtype=torch.float32, device=dev)
        gv = g_velocity.contiguous() if g_velocity is not None else torch.zeros(vs, dtype=torch.float32, device=dev)
        dd = torch.empty_like(gd)
        dv = torch.empty_like(gv)
        _sync()
        sim.tape_backward(gd.data_ptr(), gv.data_ptr(), None, dd.data_ptr(), dv.data_ptr(), None)
        return dd, dv, None, None, None


def rollout(sim, density0, velocity0, steps, checkpoint_every=8):
    """Differentiable `steps`-step rollout of `sim` (a dcgrid_amd.Simulation)."""
    return Rollout.apply(density0, velocity0, sim, steps, checkpoint_every)
