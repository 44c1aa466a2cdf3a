"""torch.autograd.Function over the device DDA solve and its adjoint gradient.

``dda_objective(rho, problem, ...)`` solves (diag(al(rho)) + A) P = E_inc for
the per-voxel densities ``rho`` (float64 CUDA tensor, x-fastest) and returns
the objective; ``backward`` runs one adjoint solve on the same device system
(the reference's calculateGradients, src/DDAModel.cpp:398-469, with exact
partials) and returns d objective / d rho.  Densities are clipped to [0, 1]
inside, as the permittivity mix is only defined there.

Objectives: "integrated" (log sum |al P|^power, ObjDDAModel.cpp:130-212) and
"targets" (mean |E|^2 over external points in nm, the field of
DDAModel::solve_E, src/DDAModel.cpp:1153-1183).
"""
from __future__ import annotations

import torch

from . import native


class _DDAObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, problem, kind, power, targets, maxit, tol):
        if not rho.is_cuda or rho.dtype != torch.float64:
            raise TypeError("rho must be a float64 CUDA tensor")
        r = rho.detach().contiguous().reshape(-1)
        torch.cuda.current_stream().synchronize()
        problem.set_density(r)
        info = problem.solve(maxit=maxit, tol=tol)
        if info["status"] != "converged":
            raise RuntimeError(f"forward solve did not converge: {info}")
        value = problem.objective(kind=kind, power=power, targets=targets)
        ctx.problem, ctx.kind, ctx.power, ctx.targets = problem, kind, power, targets
        ctx.maxit, ctx.tol, ctx.shape = maxit, tol, rho.shape
        ctx.forward_info = info
        return torch.tensor(value, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, grad_out):
        g = torch.empty(ctx.problem.n, dtype=torch.float64, device=grad_out.device)
        torch.cuda.current_stream().synchronize()
        _, _, info = ctx.problem.gradient(kind=ctx.kind, power=ctx.power, targets=ctx.targets,
                                          maxit=ctx.maxit, tol=ctx.tol, out=g)
        if info["status"] != "converged":
            raise RuntimeError(f"adjoint solve did not converge: {info}")
        return (grad_out * g).reshape(ctx.shape), None, None, None, None, None, None


def dda_objective(rho, problem, kind="integrated", power=2, targets=None, maxit=100000, tol=1e-5):
    """Differentiable objective of a native.Problem at densities rho."""
    return _DDAObjective.apply(rho, problem, kind, power, targets, maxit, tol)


def cone_filter(rho_zyx, radius):
    """Linear cone (hat) density filter of the given radius in voxels, as a 3-D
    convolution with weights max(0, radius - |offset|), normalised per voxel
    (the 3-D counterpart of the reference's 2-D cone weights,
    src/DDAModel.cpp:66-77, 580-712).  rho_zyx: (nz, ny, nx) tensor."""
    import torch.nn.functional as F
    r = int(radius)
    ax = torch.arange(-r, r + 1, dtype=rho_zyx.dtype, device=rho_zyx.device)
    dz, dy, dx = torch.meshgrid(ax, ax, ax, indexing="ij")
    w = torch.clamp(radius - torch.sqrt(dx * dx + dy * dy + dz * dz), min=0.0)
    x = rho_zyx[None, None]
    num = F.conv3d(x, w[None, None], padding=r)
    den = F.conv3d(torch.ones_like(x), w[None, None], padding=r)
    return (num / den)[0, 0]


def target_block(shape, d, size=4, height_nm=60.0):
    """size^3 target points at spacing d, centred laterally, the lowest layer
    height_nm above the top dipole layer (coordinates in nm, origin at voxel 0)."""
    nx, ny, nz = shape
    pts = []
    x0 = d * ((nx - size) / 2.0)
    y0 = d * ((ny - size) / 2.0)
    z0 = d * (nz - 1) + height_nm
    for k in range(size):
        for j in range(size):
            for i in range(size):
                pts.append((x0 + d * i, y0 + d * j, z0 + d * k))
    return pts
