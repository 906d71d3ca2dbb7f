"""TEST INFRASTRUCTURE ONLY — CPU restatement of the reference's static drivers on top
of the oracle's assembly, with dense direct solves (numpy), for small models.

  solve_linear_static(o, model)       numerics::solve_linear_static
                                      (src/numerics/linear_static.cpp:29-60)
  solve_nonlinear_static(o, model)    numerics::solve_nonlinear_static
                                      (src/numerics/nonlinear_static.cpp:197-295):
                                      run_increments + newton_iterate, no line search,
                                      default linear ramp, single point constraints

`o` is an oracle.Oracle (port or reference assembly). The reference's own drivers
need SciPP (absent), so these are the checkers of the device drivers in
calculixpp_amd/solve.py; the analytic pins (uniaxial J2 bar, elastic Newton = linear)
live in the tests.
"""
from __future__ import annotations

import numpy as np


def _expand(model, dof_eq, prescribed, uf, lam):
    u = np.zeros(3 * model.n_nodes)
    free = dof_eq >= 0
    u[free] = uf[dof_eq[free]]
    u[~free] = lam * prescribed[~free]
    return u


def _prescribed(model):
    p = np.zeros(3 * model.n_nodes)
    for n, c, v in zip(model.spc_node, model.spc_comp, model.spc_value):
        p[3 * int(n) + int(c) - 1] = v
    return p


def solve_linear_static(o, model):
    s = o.assemble_linear_static(model)
    uf = np.linalg.solve(s.dense(), s.rhs)
    return _expand(model, s.dof_eq, _prescribed(model), uf, 1.0).reshape(-1, 3)


def solve_nonlinear_static(o, model, controls=None, increment=None):
    """Returns (u (n_nodes, 3), report dict, committed+trial state (P, 26))."""
    from calculixpp_amd.solve import Incrementation, NonlinearControls
    c = controls or NonlinearControls()
    inc = increment or Incrementation()
    f_ext_full = model.external_load()
    s0 = o.assemble_linear_static(model)
    dof_eq, n_free = s0.dof_eq, s0.n_free
    free = dof_eq >= 0
    presc = _prescribed(model)
    f_ext_norm = np.linalg.norm(f_ext_full[free])
    state = np.zeros((model.n_points, 26))
    uf = np.zeros(n_free)
    rep = {"increments": 0, "iterations": 0, "cutbacks": 0}

    def newton(lam):
        nonlocal uf, state
        denom = 0.0
        for it in range(c.max_iterations):
            u = _expand(model, dof_eq, presc, uf, lam)
            sys_, f_int, st = o.assemble_material_tangent(model, u, state)
            state = st
            r = np.zeros(n_free)
            r[dof_eq[free]] = lam * f_ext_full[free] - f_int[free]
            if it == 0:
                denom = max(lam * f_ext_norm, np.linalg.norm(f_int[free]), c.eps)
            du = np.linalg.solve(sys_.dense(), r)
            uf = uf + du
            ratio = np.linalg.norm(du) / max(np.linalg.norm(uf), c.eps)
            if np.linalg.norm(r) / denom < c.force_tol and ratio < c.disp_tol:
                return it + 1
        return -1

    total = max(inc.total, c.eps)
    lam = 0.0
    dl = min(max(inc.initial / total, inc.min / total), inc.max / total)
    while lam < 1.0 - c.eps and rep["increments"] < inc.max_increments:
        step = min(1.0, lam + dl) - lam
        saved = uf.copy()
        iters = newton(lam + step)
        if iters < 0:
            if inc.direct or step * total <= inc.min + c.eps:
                break
            uf = saved
            dl = max(inc.min / total, dl * inc.cutback)
            rep["cutbacks"] += 1
            continue
        state[:, :13] = state[:, 13:]  # MaterialPoints::commit
        lam += step
        rep["increments"] += 1
        rep["iterations"] += iters
        if not inc.direct and iters <= inc.grow_below:
            dl = min(inc.max / total, dl * inc.grow)
    rep["final_load_factor"] = lam
    return _expand(model, dof_eq, presc, uf, lam).reshape(-1, 3), rep, state
