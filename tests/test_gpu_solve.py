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
    iteration counts, displacements and committed history rel <= 1e-10;
  - the recovered fields: the reference's test_hex (one C3D8 cube in uniaxial
    tension: s_xx = E delta, no lateral stress, x-face reactions -E delta) and
    test_plasticity's uniaxial end-to-end cube (mean s_zz = sy0 + H eps_p), and the
    oracle's recovery of the same displacement — rel <= 1e-10.
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
    f = gpu.solve_linear_static(m)
    u = f.displacement
    u_e = newton.solve_linear_static(port, m)
    assert rel(u, u_e) <= 1e-10
    assert u[:, 2].min() < 0  # the tip load points down
    # the recovered fields of the solved displacement match the oracle's recovery
    o = port.recover_fields(m, u_e)
    for k in ("stress", "strain", "reaction"):
        assert rel(getattr(f, k), o[k]) <= 1e-10, k
    # equilibrium: the reactions balance the tip load
    assert abs(f.reaction[:, 2].sum() + m.external_load()[2::3].sum()) <= 1e-9


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
    u_lin = gpu.solve_linear_static(m).displacement
    f_nl, rep, _ = gpu.solve_nonlinear_static(m)
    assert rep.converged and rep.increments == 1
    assert rel(f_nl.displacement, u_lin) <= 1e-10


def test_uniaxial_j2_bar_closed_form(gpu, port):
    from oracle import newton
    E, nu, sy0, H, eps = 210000.0, 0.3, 250.0, 2000.0, 0.01
    m = _bar(2, eps)
    tight = gpu.NonlinearControls(force_tol=1e-12, disp_tol=1e-12)
    f, rep, mp = gpu.solve_nonlinear_static(m, controls=tight)
    u = f.displacement
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
    f, rep, mp = gpu.solve_nonlinear_static(m)
    u = f.displacement
    u_e, rep_e, st_e = newton.solve_nonlinear_static(port, m)
    assert rep.converged and rep_e["final_load_factor"] == 1.0
    assert (rep.increments, rep.iterations, rep.cutbacks) == \
        (rep_e["increments"], rep_e["iterations"], rep_e["cutbacks"])
    assert rel(u, u_e) <= 1e-10
    st = mp.committed.cpu().numpy()
    assert (st[6] > 0).mean() > 0.1  # a good part of the points yielded
    for k in range(13):
        assert rel(st[k], st_e[:, k]) <= 1e-10, k


def _cube(uniaxial_x: bool, delta: float, j2=None):
    """The single C3D8 unit cube of the reference's test_hex (make_cube: x = 0 face
    held in x, x = 1 face at u_x = delta, rigid modes removed at nodes 1, 4, 5) or of
    test_plasticity (uniaxial_cube: symmetry on the three back faces, top at u_z)."""
    from calculixpp_amd import _lib
    from calculixpp_amd.model import Model
    x = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    conn = np.arange(8, dtype=np.int32)[None, :]
    if uniaxial_x:
        spc = [(n, 1, 0.0) for n in (1, 4, 5, 8)] + [(1, 2, 0.0), (1, 3, 0.0), (4, 3, 0.0),
                                                     (5, 2, 0.0)]
        spc += [(n, 1, delta) for n in (2, 3, 6, 7)]
    else:
        spc = [(n, 1, 0.0) for n in (1, 4, 5, 8)] + [(n, 2, 0.0) for n in (1, 2, 5, 6)]
        spc += [(n, 3, 0.0) for n in (1, 2, 3, 4)] + [(n, 3, delta) for n in (5, 6, 7, 8)]
    kw = {} if j2 is None else dict(j2_id=np.zeros(1, np.int32), j2=[j2])
    return Model(x, [(_lib.C3D8, conn)], 210000.0, 0.3,
                 spc_node=np.array([s[0] - 1 for s in spc]),
                 spc_comp=np.array([s[1] for s in spc]),
                 spc_value=np.array([s[2] for s in spc]), **kw)


def test_hex_uniaxial_fields(gpu, port):
    """tests/test_hex.cpp:58-81 on the device: homogeneous uniaxial solution."""
    from oracle import newton
    E, nu, delta = 210000.0, 0.3, 0.01
    m = _cube(True, delta)
    f = gpu.solve_linear_static(m)
    i7 = 6
    assert abs(f.displacement[i7, 0] - delta) <= 1e-10
    assert abs(f.displacement[i7, 1] + nu * delta) <= 1e-10
    assert abs(f.stress[i7, 0] - E * delta) <= 1e-9 * E * delta
    assert abs(f.stress[i7, 1]) <= 1e-9 * E * delta and abs(f.stress[i7, 2]) <= 1e-9 * E * delta
    assert abs(f.strain[i7, 0] - delta) <= 1e-12 and abs(f.strain[i7, 1] + nu * delta) <= 1e-12
    rf0 = f.reaction[[0, 3, 4, 7], 0].sum()
    assert abs(rf0 + E * delta) <= 1e-9 * E * delta
    o = port.recover_fields(m, newton.solve_linear_static(port, m))
    for k in ("stress", "strain", "reaction"):
        assert rel(getattr(f, k), o[k]) <= 1e-10, k


def test_plastic_cube_fields(gpu, port):
    """tests/test_plasticity.cpp:162-197 on the device: the recovered axial stress of
    a J2 cube follows sy0 + H eps_p (material-point recovery)."""
    from calculixpp_amd.model import J2Params
    from oracle import newton
    from oracle.oracle import aos_state
    E, H, sy0, uz = 210000.0, 8000.0, 800.0, 0.01
    m = _cube(False, uz, J2Params(E, 0.3, [0.0, 0.02], [800.0, 960.0]))
    tight = gpu.NonlinearControls(force_tol=1e-12, disp_tol=1e-12)
    f, rep, mp = gpu.solve_nonlinear_static(m, controls=tight)
    assert rep.converged
    eps_p = (E * uz - sy0) / (E + H)
    sigma = sy0 + H * eps_p
    assert abs(f.stress[:, 2].mean() - sigma) <= 1e-9 * sigma
    assert np.max(np.abs(f.stress[:, :2])) <= 1e-8 * sigma  # uniaxial
    u_e, _, st_e = newton.solve_nonlinear_static(port, m, controls=tight)
    o = port.recover_fields(m, u_e, st_e)
    for k in ("stress", "strain", "reaction"):
        assert rel(getattr(f, k), o[k]) <= 1e-10, k
