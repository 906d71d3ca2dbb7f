"""Displacement solves on the device: linear static and the Newton-Raphson driver.

  solve_linear_static(model)       numerics::solve_linear_static (src/numerics/
                                   linear_static.cpp:29-50): assemble K_ff, rhs; solve;
                                   expand to the nodal displacement field; recover
                                   stress, strain and reactions
  solve_nonlinear_static(model)    numerics::solve_nonlinear_static (src/numerics/
                                   nonlinear_static.cpp:266-295): load factor 0 -> 1
                                   with automatic incrementation and cutback
                                   (run_increments, :197-233), a Newton loop per
                                   increment (newton_iterate, :159-186) whose
                                   assemble + return map + residual is one HIP graph,
                                   a commit of the material points per increment,
                                   and the field recovery of :276-288

The sparse solves run on the device (cxg_pcg, Jacobi-preconditioned CG) in place of
SciPP (ComputeBackend::solve_sparse). Both drivers return the reference's
StaticFields (calculixpp_amd.recover, cxg_recover_fields). Loads and prescribed values
follow the default linear ramp (no *AMPLITUDE), the models carry single point
constraints only, and the Newton loop takes no line search (NonlinearOptions
defaults).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import fem
from .model import Model
from .recover import FieldRecovery, StaticFields


@dataclass
class NonlinearControls:
    """core/solution_control.hpp:11-20 defaults."""
    force_tol: float = 5e-3
    disp_tol: float = 1e-3
    max_iterations: int = 16
    eps: float = 1e-30


@dataclass
class Incrementation:
    """core/solution_control.hpp:25-35 defaults."""
    initial: float = 1.0
    total: float = 1.0
    min: float = 1e-5
    max: float = 1.0
    direct: bool = False
    cutback: float = 0.25
    grow: float = 1.5
    grow_below: int = 5
    max_increments: int = 1000


def pcg(ctx: fem.Context, row_ptr: torch.Tensor, col_idx: torch.Tensor, vals: torch.Tensor,
        b: torch.Tensor, x: Optional[torch.Tensor] = None, tol: float = 1e-13,
        max_iter: Optional[int] = None):
    """Solve A x = b (CSR on the device) by cxg_pcg; returns (x, iterations, rel residual)."""
    n = int(b.numel())
    x = torch.zeros_like(b) if x is None else x
    it, rr = C.c_int32(), C.c_double()
    ctx.check(ctx.lib.cxg_pcg(ctx.h, n, fem._ptr(row_ptr), fem._ptr(col_idx), fem._ptr(vals),
                              fem._ptr(b), fem._ptr(x), float(tol),
                              int(max_iter if max_iter is not None else max(100, 20 * n)),
                              C.byref(it), C.byref(rr)))
    return x, it.value, rr.value


def _expand(n_nodes: int, free_idx: torch.Tensor, spc_idx: torch.Tensor, spc_val: torch.Tensor,
            uf: torch.Tensor, lam: float, out: torch.Tensor) -> torch.Tensor:
    """expand_displacement (nonlinear_static.cpp:42-60) for single point constraints:
    free DOFs from uf, prescribed DOFs at lam * value."""
    out.zero_()
    out[free_idx] = uf
    if spc_idx.numel():
        out[spc_idx] = lam * spc_val
    return out


def solve_linear_static(model: Model, tol: float = 1e-13, device=None) -> StaticFields:
    """Displacements, nodal stress and strain, and reactions of a linear static step."""
    ctx = fem.context(device)
    a = fem.Assembler(model, ctx)
    vals = a.new_values()
    a.assemble(vals)
    rhs = a.rhs()
    x, _, _ = pcg(ctx, a.pattern.row_ptr, a.pattern.col_idx, vals[: a.pattern.nnz], rhs, tol=tol)
    dev = ctx.device
    free = torch.from_numpy(np.nonzero(a.dof_eq >= 0)[0]).to(dev)
    spc = torch.from_numpy(np.nonzero(a.dof_eq < 0)[0]).to(dev)
    u = torch.empty(3 * model.n_nodes, dtype=torch.float64, device=dev)
    _expand(model.n_nodes, free, spc, torch.from_numpy(a.prescribed[a.dof_eq < 0]).to(dev),
            x, 1.0, u)
    return _fields(model, ctx, u)


def _fields(model: Model, ctx: fem.Context, u: torch.Tensor,
            mp: Optional[fem.MaterialPoints] = None) -> StaticFields:
    """recover_fields (linear_static.cpp:48; nonlinear_static.cpp:284-288: the
    material-point form for models with a nonlinear material)."""
    o = FieldRecovery(model, ctx).run(u.reshape(-1, 3), mp if model.j2 else None)
    return StaticFields(u.reshape(-1, 3).cpu().numpy(), o["stress"].cpu().numpy(),
                        o["strain"].cpu().numpy(), o["reaction"].cpu().numpy())


@dataclass
class NonlinearReport:
    """numerics/nonlinear_static.hpp:33-39."""
    increments: int = 0
    iterations: int = 0
    cutbacks: int = 0
    converged: bool = False
    final_load_factor: float = 0.0


class _Newton:
    """The per-step device state of solve_nonlinear_static (NewtonState,
    nonlinear_static.cpp:72-84)."""

    def __init__(self, model: Model, ctx: fem.Context, cg_tol: float):
        self.model, self.ctx, self.cg_tol = model, ctx, cg_tol
        dev = ctx.device
        self.asm = fem.TangentAssembler(model, ctx)
        dof_eq = self.asm.dof_eq
        self.n_free = self.asm.n_free
        self.free = torch.from_numpy(np.nonzero(dof_eq >= 0)[0]).to(dev)
        self.spc = torch.from_numpy(np.nonzero(dof_eq < 0)[0]).to(dev)
        self.spc_val = torch.from_numpy(self.asm.prescribed[dof_eq < 0].copy()).to(dev)
        self.f_ext_full = torch.from_numpy(model.external_load()).to(dev)
        self.f_ext_norm = float(torch.linalg.norm(self.f_ext_full[self.free]))
        self.mp = fem.make_material_points(model, dev)  # virgin history (assembly.cpp:357-383)
        n3 = 3 * model.n_nodes
        self.u = torch.zeros(n3, dtype=torch.float64, device=dev)
        self.f_ext = torch.zeros(n3, dtype=torch.float64, device=dev)
        self.vals = self.asm.new_values()
        self.f_int = torch.empty(n3, dtype=torch.float64, device=dev)
        self.rhs = torch.empty(max(self.n_free, 1), dtype=torch.float64, device=dev)
        self.uf = torch.zeros(self.n_free, dtype=torch.float64, device=dev)
        self.graph = self.asm.capture(self.u, self.mp, self.vals, self.f_int, self.f_ext, self.rhs)
        self.cg_iterations = 0

    def assemble_at(self, lam: float):
        """assemble_at (nonlinear_static.cpp:91-113): K_t in vals, r = lam f_ext,f - f_int,f
        in rhs; returns ||f_int,f||."""
        _expand(self.model.n_nodes, self.free, self.spc, self.spc_val, self.uf, lam, self.u)
        torch.mul(self.f_ext_full, lam, out=self.f_ext)
        self.graph.launch()
        self.ctx.check_kernels()
        return float(torch.linalg.norm(self.f_int[self.free]))

    def iterate(self, c: NonlinearControls, lam: float) -> int:
        """newton_iterate (nonlinear_static.cpp:159-186), no line search."""
        denom = 0.0
        p = self.asm.pattern
        for it in range(c.max_iterations):
            f_int_norm = self.assemble_at(lam)
            r = self.rhs[: self.n_free]
            if it == 0:
                denom = max(lam * self.f_ext_norm, f_int_norm, c.eps)
            du, cg_it, _ = pcg(self.ctx, p.row_ptr, p.col_idx, self.vals[: p.nnz], r,
                               tol=self.cg_tol)
            self.cg_iterations += cg_it
            self.uf += du
            du_norm = float(torch.linalg.norm(du))
            u_norm = float(torch.linalg.norm(self.uf))
            disp_ratio = du_norm / max(u_norm, c.eps)
            if float(torch.linalg.norm(r)) / denom < c.force_tol and disp_ratio < c.disp_tol:
                return it + 1
        return -1


def solve_nonlinear_static(model: Model, controls: Optional[NonlinearControls] = None,
                           increment: Optional[Incrementation] = None, cg_tol: float = 1e-13,
                           device=None):
    """Returns (StaticFields, NonlinearReport, MaterialPoints with the committed
    history)."""
    c = controls or NonlinearControls()
    inc = increment or Incrementation()
    st = _Newton(model, fem.context(device), cg_tol)
    rep = NonlinearReport()
    # run_increments (nonlinear_static.cpp:197-233); no *TIME POINTS
    total = max(inc.total, c.eps)
    lam = 0.0
    dl = min(max(inc.initial / total, inc.min / total), inc.max / total)
    while lam < 1.0 - c.eps and rep.increments < inc.max_increments:
        step = min(1.0, lam + dl) - lam
        saved = st.uf.clone()
        iters = st.iterate(c, lam + step)
        if iters < 0:
            if inc.direct or step * total <= inc.min + c.eps:
                break
            st.uf.copy_(saved)
            dl = max(inc.min / total, dl * inc.cutback)
            rep.cutbacks += 1
            continue
        st.mp.commit()
        lam += step
        rep.increments += 1
        rep.iterations += iters
        if not inc.direct and iters <= inc.grow_below:
            dl = min(inc.max / total, dl * inc.grow)
    rep.final_load_factor = lam
    rep.converged = lam >= 1.0 - c.eps
    u = _expand(model.n_nodes, st.free, st.spc, st.spc_val, st.uf, lam,
                torch.empty_like(st.u))
    del st.graph
    return _fields(model, st.ctx, u, st.mp), rep, st.mp
