"""Differentiable velocity field: torch.autograd.Function over the HIP
velocity chain and its hand-written adjoint (gpk_velocity_field /
gpk_velocity_adjoint).

    v = VelocityField.apply(theta, grid)

theta is a CPU float64 tensor (eps4, eps6, M0); v is a device tensor of the
grid's shape and real type holding the projected interface velocity
v = P M (dG - sigma(n; eps4, eps6) kappa) on band cells and 0 elsewhere
(src/gpk_physics.cpp:67-193; the eps6 cos 6 theta term is an extension).
Backward returns dL/dtheta; the labels, normals and curvature are not
differentiated (the flips are discrete) and are recomputed on the device in
both passes rather than stored.
"""
import torch


def _device(grid):
    return torch.device("cuda", grid.device)


class VelocityField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, grid):
        th = [float(t) for t in theta.detach().cpu().reshape(3)]
        dtype = torch.float64 if grid.precision == "fp64" else torch.float32
        v = torch.empty(grid.shape, dtype=dtype, device=_device(grid))
        # the library runs on its own stream: torch's queued work on v's
        # memory must be finished first
        torch.cuda.current_stream(v.device).synchronize()
        grid.velocity_field_ptr(th, v.data_ptr())
        ctx.grid = grid
        ctx.theta = th
        ctx.dtype = dtype
        ctx.theta_dtype = theta.dtype
        return v

    @staticmethod
    def backward(ctx, grad_v):
        g = grad_v.detach().to(ctx.dtype).contiguous()
        torch.cuda.current_stream(g.device).synchronize()
        grad = ctx.grid.velocity_adjoint_ptr(ctx.theta, g.data_ptr())
        return torch.tensor(grad, dtype=ctx.theta_dtype), None


def velocity_field(theta, grid):
    """Projected interface velocity as a differentiable function of theta."""
    return VelocityField.apply(theta, grid)
