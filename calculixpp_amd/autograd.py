"""r = K(x, rho) u as a differentiable torch operation (the adjoint path).

Forward: the matrix-free internal force r = sum_e Ke(x, E_e) ue (libcxg K5,
fem::internal_force semantics, src/fem/stress.cpp:125-145) with the SIMP law
E_e = E0 * rho_e^p. Backward: the hand-written vector-Jacobian product kernel (K6,
cxg_vjp), which recomputes B at the Gauss points instead of storing it:
  dx  = d(w.r)/dx   (shape derivative of the element stiffness)
  drho = d(w.r)/dE * p E0 rho^(p-1)
  du  = K w
The reference offers only coordinate design variables, by central differences of the
re-assembled reduced operator (src/numerics/sensitivity.cpp:52-108); density and
displacement cotangents are new here.

    mesh = ResidualMesh(model)           # uploads connectivity once, colours it
    r = residual(mesh, x, rho, u)        # differentiable in x, rho and u
"""
from __future__ import annotations

from typing import Optional

import torch

from . import fem
from .model import Model


class ResidualMesh:
    """Connectivity and element families of a model, resident on the device. The
    node coordinates, densities and displacements are the differentiable inputs."""

    def __init__(self, model: Model, E0: float = 210000.0, penal: float = 3.0,
                 ctx: Optional[fem.Context] = None):
        self.mesh = fem.DeviceMesh(model, ctx)
        self.ctx = self.mesh.ctx
        self.n_nodes, self.n_elem = model.n_nodes, model.n_elem
        self.nu = self.mesh.nu
        self.E0, self.penal = float(E0), float(penal)

    def forward_raw(self, x: torch.Tensor, E: torch.Tensor, u: torch.Tensor,
                    out: Optional[torch.Tensor] = None) -> torch.Tensor:
        r = out if out is not None else torch.empty_like(u)
        self.ctx.check(self.ctx.lib.cxg_internal_force(
            self.ctx.h, self.mesh.h, fem._ptr(x), fem._ptr(E), fem._ptr(self.nu), fem._ptr(u),
            fem._ptr(r)))
        return r

    def vjp_raw(self, x: torch.Tensor, E: torch.Tensor, u: torch.Tensor, w: torch.Tensor,
                dx: Optional[torch.Tensor] = None, dE: Optional[torch.Tensor] = None,
                du: Optional[torch.Tensor] = None):
        dx = dx if dx is not None else torch.empty_like(x)
        du = du if du is not None else torch.empty_like(u)
        dE = dE if dE is not None else torch.empty_like(E)
        self.ctx.check(self.ctx.lib.cxg_vjp(
            self.ctx.h, self.mesh.h, fem._ptr(x), fem._ptr(E), fem._ptr(self.nu), fem._ptr(u),
            fem._ptr(w), fem._ptr(dx), fem._ptr(dE), fem._ptr(du)))
        return dx, dE, du


def _prep(t: torch.Tensor, shape) -> torch.Tensor:
    if t.dtype != torch.float64:
        raise TypeError("the assembly path computes in FP64; pass float64 tensors")
    return t.reshape(shape).contiguous()


class ResidualFunction(torch.autograd.Function):
    """torch.autograd.Function for r = K(x, rho) u (see module docstring)."""

    @staticmethod
    def forward(ctx, x, rho, u, mesh: ResidualMesh):
        x = _prep(x, (mesh.n_nodes, 3))
        u = _prep(u, (mesh.n_nodes, 3))
        rho = _prep(rho, (mesh.n_elem,))
        E = mesh.E0 * rho ** mesh.penal
        r = mesh.forward_raw(x, E, u)
        ctx.save_for_backward(x, rho, u, E)
        ctx.mesh = mesh
        return r

    @staticmethod
    def backward(ctx, grad_r):
        x, rho, u, E = ctx.saved_tensors
        mesh: ResidualMesh = ctx.mesh
        w = _prep(grad_r, (mesh.n_nodes, 3))
        dx, dE, du = mesh.vjp_raw(x, E, u, w)
        drho = dE * (mesh.penal * mesh.E0) * rho ** (mesh.penal - 1.0)
        return dx, drho, du, None


def residual(mesh: ResidualMesh, x: torch.Tensor, rho: torch.Tensor, u: torch.Tensor) -> torch.Tensor:
    """r = K(x, rho) u, differentiable in x, rho and u; shape (n_nodes, 3)."""
    return ResidualFunction.apply(x, rho, u, mesh)
