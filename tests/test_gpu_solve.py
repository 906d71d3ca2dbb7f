"""Displacement solves on the device (calculixpp_amd.solve: cxg_pcg + the Newton driver).

Pinned by:
  - config 1: displacements against a dense direct solve of the oracle's K_ff, rhs
    (SciPP is absent, so its spsolve output stays unpinned) — rel <= 1e-10;
  - an elastic model through the Newton driver equals the linear solve (the
    reference's test_nonlinear: NL = linear) — rel <= 1e-10;
  - a uniaxial J2 bar in closed form: eps_p = (E eps - sigma_y0) / (E + H) at every
    Gauss point, lateral strain -nu sigma/E - eps_p/2 — rel <= 1e-10;
  - a J2 block (clamped end, prescribed tip displacement, loads) against the oracle's
    CPU restatement of the same driver (oracle/newton.py): equal increment and
    iteration counts, displacements and committed history rel <= 1e-10.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from calculixpp_amd import solve
    return solve


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_config1_displacements(gpu, port):
    from calculixpp_amd import meshgen
    from oracle import newton
    m = meshgen.cantilever()
    u = gpu.solve_linear_static(m)
    u_e = newton.solve_linear_static(port, m)
    assert rel(u, u_e) <= 1e-10
    assert u[:, 2].min() < 0  # the tip load points down


def test_pcg_is_deterministic(gpu):
    import torch
    from calculixpp_amd import fem, meshgen
    m = meshgen.cantilever()
    a = fem.Assembler(m)
    v = a.new_values()
    a.assemble(v)
    b = a.rhs()
    x1, it1, r1 = gpu.pcg(a.ctx, a.pattern.row_ptr, a.pattern.col_idx, v[: a.pattern.nnz], b)
    x2, it2, r2 = gpu.pcg(a.ctx, a.pattern.row_ptr, a.pattern.col_idx, v[: a.pattern.nnz], b)
    assert it1 == it2 and r1 <= 1e-13 and torch.equal(x1, x2)


def _bar(n=2, delta=0.01, j2=True):
    """Unit cube of n^3 C3D8, uniaxial in z: uz = 0 on z = 0, uz = delta on z = 1, the
    rigid modes removed at two bottom corners only, so the stress state is uniaxial."""
    from calculixpp_amd import _lib, meshgen
    from calculixpp_amd.model import J2Params, Model
    c, conn = meshgen.hex_grid(n, n, n, (1.0, 1.0, 1.0), _lib.C3D8)
    bot = np.nonzero(np.abs(c[:, 2]) < 1e-12)[0]
    top = np.nonzero(np.abs(c[:, 2] - 1.0) < 1e-12)[0]
    o = int(np.nonzero(np.all(np.abs(c) < 1e-12, axis=1))[0][0])
    x1 = int(np.nonzero(np.all(np.abs(c - [1.0, 0.0, 0.0]) < 1e-12, axis=1))[0][0])
    spc_node = np.concatenate([bot, top, [o, o, x1]])
    spc_comp = np.concatenate([np.full(len(bot), 3), np.full(len(top), 3), [1, 2, 2]])
    spc_val = np.concatenate([np.zeros(len(bot)), np.full(len(top), delta), [0.0, 0.0, 0.0]])
    kw = {}
    if j2:
        kw = dict(j2_id=np.zeros(len(conn), np.int32),
                  j2=[J2Params(210000.0, 0.3, [0.0, 1.0], [250.0, 2250.0])])
    return Model(c, [(_lib.C3D8, conn)], 210000.0, 0.3, spc_node=spc_node, spc_comp=spc_comp,
                 spc_value=spc_val, **kw)


def test_elastic_newton_equals_linear(gpu):
    from calculixpp_amd import meshgen
    m = meshgen.cantilever()
    u_lin = gpu.solve_linear_static(m)
    u_nl, rep, _ = gpu.solve_nonlinear_static(m)
    assert rep.converged and rep.increments == 1
    assert rel(u_nl, u_lin) <= 1e-10


def test_uniaxial_j2_bar_closed_form(gpu, port):
    from oracle import newton
    E, nu, sy0, H, eps = 210000.0, 0.3, 250.0, 2000.0, 0.01
    m = _bar(2, eps)
    tight = gpu.NonlinearControls(force_tol=1e-12, disp_tol=1e-12)
    u, rep, mp = gpu.solve_nonlinear_static(m, controls=tight)
    assert rep.converged
    ep = (E * eps - sy0) / (E + H)
    sig = sy0 + H * ep
    lat = -nu * sig / E - 0.5 * ep
    st = mp.committed.cpu().numpy()
    assert np.max(np.abs(st[6] - ep)) <= 1e-10 * ep
    assert np.max(np.abs(st[2] - ep)) <= 1e-10 * ep          # axial plastic strain
    assert np.max(np.abs(st[0] + 0.5 * ep)) <= 1e-10 * ep    # lateral, plastic incompressibility
    x1 = np.abs(m.coords[:, 0] - 1.0) < 1e-12
    assert np.max(np.abs(u[x1, 0] - lat)) <= 1e-10 * abs(lat)
    u_e, rep_e, st_e = newton.solve_nonlinear_static(port, m, controls=tight)
    assert rel(u, u_e) <= 1e-10


def test_j2_block_newton_matches_oracle(gpu, port):
    from calculixpp_amd import _lib, meshgen
    from calculixpp_amd.model import J2Params, Model
    from oracle import newton
    c, conn = meshgen.hex_grid(4, 2, 2, (2.0, 0.5, 0.5), _lib.C3D8)
    c = meshgen.jitter(c, 0.25, 0.1, 19, (2.0, 0.5, 0.5))
    x0 = np.nonzero(c[:, 0] < 1e-12)[0]
    x1 = np.nonzero(np.abs(c[:, 0] - 2.0) < 1e-12)[0]
    spc_node = np.concatenate([np.repeat(x0, 3), x1])
    spc_comp = np.concatenate([np.tile([1, 2, 3], len(x0)), np.full(len(x1), 3)])
    spc_val = np.concatenate([np.zeros(3 * len(x0)), np.full(len(x1), -0.004)])
    m = Model(c, [(_lib.C3D8, conn)], 210000.0, 0.3, spc_node=spc_node, spc_comp=spc_comp,
              spc_value=spc_val, cload_node=x1[:2], cload_comp=np.array([2, 2]),
              cload_value=np.array([5.0, 5.0]),
              j2_id=np.zeros(len(conn), np.int32),
              j2=[J2Params(210000.0, 0.3, [0.0, 1.0], [250.0, 2250.0])])
    u, rep, mp = gpu.solve_nonlinear_static(m)
    u_e, rep_e, st_e = newton.solve_nonlinear_static(port, m)
    assert rep.converged and rep_e["final_load_factor"] == 1.0
    assert (rep.increments, rep.iterations, rep.cutbacks) == \
        (rep_e["increments"], rep_e["iterations"], rep_e["cutbacks"])
    assert rel(u, u_e) <= 1e-10
    st = mp.committed.cpu().numpy()
    assert (st[6] > 0).mean() > 0.1  # a good part of the points yielded
    for k in range(13):
        assert rel(st[k], st_e[:, k]) <= 1e-10, k
