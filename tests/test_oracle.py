"""Pins of the oracle (CPU restatement, oracle/cxg_oracle.cpp).

Three kinds of checks, so that the oracle is trusted for what it is used for:
  1. golden fixtures written from the reference core (tests/golden/*.npz) —
     bit-exact;
  2. the reference core itself (oracle/_ref, when built here) on fresh seeds —
     bit-exact;
  3. what the mathematics fixes: partition of unity, gradient consistency, exact
     volumes, symmetry and rigid-body null space, the constant-strain patch test,
     dense brute-force assembly, the J2 closed forms and a finite-difference
     tangent, and central finite differences of the adjoint.
"""
import os

import numpy as np
import pytest

from calculixpp_amd import _lib, meshgen
from calculixpp_amd.model import J2Params, Model

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FAMS = {"C3D8": _lib.C3D8, "C3D20": _lib.C3D20, "C3D20R": _lib.C3D20R, "C3D6": _lib.C3D6}
NN = {"C3D8": 8, "C3D20": 20, "C3D20R": 20, "C3D6": 6}


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# ---- 1. golden fixtures from the reference core -----------------------------------

def test_golden_cantilever(port):
    g = gold("cantilever_c3d8")
    s = port.assemble_linear_static(meshgen.cantilever())
    rp, ci, v = s.csr()
    assert s.n_free == 1500 and len(v) == 88218
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(ci, g["col_idx"])
    assert np.array_equal(v, g["vals"]) and np.array_equal(s.rhs, g["rhs"])
    assert np.array_equal(s.dof_eq, g["dof_eq"])
    assert int((v == 0.0).sum()) == 1148  # exact-zero sums kept as entries


@pytest.mark.parametrize("name", list(FAMS))
def test_golden_single_elements(port, name):
    g = gold("elements")
    x, (E, nu) = g[name + "_coords"], g[name + "_Enu"]
    assert np.array_equal(port.element_stiffness(FAMS[name], x, E, nu), g[name + "_Ke"])
    _, fe, _ = port.element_tangent_force(FAMS[name], x, g[name + "_ue"], E, nu)
    assert np.array_equal(fe, g[name + "_fe"])


@pytest.mark.parametrize("name", list(FAMS))
def test_golden_blocks(port, name):
    g = gold("blocks_2x2x2")
    m = meshgen.block(2, FAMS[name], seed_geo=21, seed_E=22)
    s = port.assemble_linear_static(m)
    rp, ci, v = s.csr()
    assert np.array_equal(rp, g[name + "_row_ptr"]) and np.array_equal(ci, g[name + "_col_idx"])
    assert np.array_equal(v, g[name + "_vals"])
    assert np.array_equal(port.internal_force(m, g[name + "_u"]), g[name + "_fint"])


def _params(g, key):
    E, nu, hard, hk = g[key + "_params"]
    t = g[key + "_table"]
    return J2Params(E, nu, list(t[0]), list(t[1]), int(hard), hk)


@pytest.mark.parametrize("key", ["iso", "kin", "comb"])
def test_golden_j2_points(port, key):
    g = gold("j2_points")
    p = _params(g, key)
    for i in range(len(g[key + "_strain"])):
        sig, D, st = port.j2_evaluate(p, g[key + "_strain"][i], g[key + "_state_in"][i])
        assert np.array_equal(sig, g[key + "_stress"][i])
        assert np.array_equal(D, g[key + "_D"][i])
        assert np.array_equal(st, g[key + "_state_out"][i])


def test_golden_j2_mesh(port):
    from oracle.oracle import aos_state
    g = gold("j2_mesh_c3d8")
    m = meshgen.j2_block(3, 3, 2, (1.5, 1.5, 1.0))
    s, fi, st = port.assemble_material_tangent(m, g["u"], aos_state(g["history"]))
    rp, ci, v = s.csr()
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(ci, g["col_idx"])
    assert np.array_equal(v, g["vals"]) and np.array_equal(fi, g["f_int"])
    assert np.array_equal(st, g["state_out"])


# ---- 2. the reference core on fresh seeds ------------------------------------------

@pytest.mark.parametrize("name", list(FAMS))
def test_reference_bitwise_blocks(port, ref, name):
    m = meshgen.block(3, FAMS[name], seed_geo=101, seed_E=102)
    a, b = port.assemble_linear_static(m), ref.assemble_linear_static(m)
    for x, y in zip(a.csr(), b.csr()):
        assert np.array_equal(x, y)
    u = meshgen.smooth_field(m.coords, 103)
    assert np.array_equal(port.internal_force(m, u), ref.internal_force(m, u))


def test_reference_bitwise_j2(port, ref):
    from oracle.oracle import aos_state
    m = meshgen.j2_block(4, 3, 3, (1.0, 0.75, 0.75))
    h = meshgen.seeded_history(m.n_points, 104)
    u = meshgen.smooth_field(m.coords, 105, 3e-3)
    sa, fa, sta = port.assemble_material_tangent(m, u, aos_state(h))
    sb, fb, stb = ref.assemble_material_tangent(m, u, aos_state(h))
    for x, y in zip(sa.csr(), sb.csr()):
        assert np.array_equal(x, y)
    assert np.array_equal(fa, fb) and np.array_equal(sta, stb)


def test_reference_bitwise_prescribed_and_loads(port, ref):
    coords, conn = meshgen.hex_grid(3, 2, 2, (1.5, 1.0, 1.0), _lib.C3D20R)
    rng = np.random.default_rng(106)
    x0 = np.nonzero(coords[:, 0] < 1e-12)[0]
    m = Model(coords, [(_lib.C3D20R, conn)], rng.uniform(1e5, 2e5, len(conn)), 0.3,
              spc_node=np.repeat(x0, 2), spc_comp=np.tile([1, 3], len(x0)),
              spc_value=rng.uniform(-1e-3, 1e-3, 2 * len(x0)),
              cload_node=rng.choice(len(coords), 5), cload_comp=rng.integers(1, 4, 5),
              cload_value=rng.normal(size=5))
    a, b = port.assemble_linear_static(m), ref.assemble_linear_static(m)
    for x, y in zip(a.csr(), b.csr()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.rhs, b.rhs)


# ---- 3. what the mathematics fixes -------------------------------------------------

def _shape(port, t, p):
    import ctypes as C
    N, dN = np.zeros(20), np.zeros(60)
    f = port.L.orc_shape
    f.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    n = f(t, *p, N.ctypes.data_as(C.c_void_p), dN.ctypes.data_as(C.c_void_p))
    return N[:n], dN[: 3 * n].reshape(n, 3)


@pytest.mark.parametrize("name", list(FAMS))
def test_partition_of_unity_and_derivative_fd(port, name):
    rng = np.random.default_rng(1)
    for _ in range(5):
        p = rng.uniform(0.05, 0.3, 3) if name == "C3D6" else rng.uniform(-0.9, 0.9, 3)
        N, dN = _shape(port, FAMS[name], p)
        assert abs(N.sum() - 1.0) < 1e-13
        assert np.abs(dN.sum(axis=0)).max() < 1e-13
        h = 1e-6
        for j in range(3):  # dN is the derivative of N (central difference)
            e = np.zeros(3)
            e[j] = h
            Np, _ = _shape(port, FAMS[name], p + e)
            Nm, _ = _shape(port, FAMS[name], p - e)
            assert np.abs((Np - Nm) / (2 * h) - dN[:, j]).max() < 1e-8


def _affine_element(name, seed):
    rng = np.random.default_rng(seed)
    t = FAMS[name]
    coords, conn = meshgen.hex_grid(1, 1, 1, (2.0, 2.0, 2.0), t)
    ref_x = coords[conn[0]] - 1.0  # reference cube [-1,1]^3
    A = np.eye(3) + rng.uniform(-0.3, 0.3, (3, 3))
    b = rng.normal(size=3)
    return ref_x @ A.T + b, A


@pytest.mark.parametrize("name", list(FAMS))
def test_gradients_consistency_and_volume(port, name):
    import ctypes as C
    x, A = _affine_element(name, 3)
    x = x + np.random.default_rng(4).uniform(-0.02, 0.02, x.shape)  # no longer affine
    f = port.L.orc_gradients
    f.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
    for q in range(port.L.orc_num_gauss(FAMS[name])):
        g, det = np.zeros((NN[name], 3)), C.c_double()
        xs = np.ascontiguousarray(x)
        assert f(FAMS[name], q, xs.ctypes.data_as(C.c_void_p), g.ctypes.data_as(C.c_void_p),
                 C.byref(det), None, 0) == 0
        assert det.value > 0
        assert np.abs(g.sum(axis=0)).max() < 1e-12
        assert np.abs(g.T @ x - np.eye(3)).max() < 1e-12
    # exact volume of an affine image: |det A| times the reference volume
    xa, A = _affine_element(name, 5)
    vol = 0.0
    rule = np.zeros(27 * 4)
    nq = port.L.orc_gauss_rule(FAMS[name], rule.ctypes.data_as(C.c_void_p))
    for q in range(nq):
        g, det = np.zeros((NN[name], 3)), C.c_double()
        xs = np.ascontiguousarray(xa)
        f(FAMS[name], q, xs.ctypes.data_as(C.c_void_p), g.ctypes.data_as(C.c_void_p),
          C.byref(det), None, 0)
        vol += det.value * rule[4 * q + 3]
    v_ref = 4.0 if name == "C3D6" else 8.0  # wedge = half of the [-1,1]^3 cube
    assert abs(vol - v_ref * abs(np.linalg.det(A))) < 1e-12 * vol


def _rigid_modes(x):
    modes = []
    for d in range(3):
        m = np.zeros_like(x)
        m[:, d] = 1.0
        modes.append(m.ravel())
    for axis in np.eye(3):
        modes.append(np.cross(axis, x).ravel())
    return modes


@pytest.mark.parametrize("name", list(FAMS))
def test_stiffness_symmetry_and_rigid_modes(port, name):
    x, _ = _affine_element(name, 6)
    x = x + np.random.default_rng(7).uniform(-0.05, 0.05, x.shape)
    K = port.element_stiffness(FAMS[name], x, 2.1e5, 0.3)
    assert np.array_equal(K, K.T)
    for r in _rigid_modes(x):
        assert np.linalg.norm(K @ r) < 1e-10 * np.linalg.norm(K) * np.linalg.norm(r)
    w = np.linalg.eigvalsh(K)
    # full rules: exactly the 6 rigid-body modes. A lone C3D20R under its reduced
    # 2x2x2 rule has 6 spurious zero-energy modes besides; the reference's C3D6 rule
    # (triangle centroid x 2 points, element.cpp:68-71) leaves one.
    expected = {"C3D8": 6, "C3D20": 6, "C3D20R": 12, "C3D6": 7}[name]
    assert np.sum(w < 1e-8 * w.max()) == expected


@pytest.mark.parametrize("name", list(FAMS))
def test_patch_test_constant_strain(port, name):
    # the reduced C3D20R rule integrates exactly only on undistorted elements
    jittered = meshgen.block(3, FAMS[name], seed_geo=8,
                             jitter_frac=0.0 if name == "C3D20R" else 0.1)
    m = Model(jittered.coords, jittered.blocks, 2.0e5, 0.3)  # one material: sigma uniform
    A = np.random.default_rng(9).normal(size=(3, 3)) * 1e-3
    u = m.coords @ A.T
    f = port.internal_force(m, u).reshape(-1, 3)
    interior = np.all((m.coords > 1e-9) & (m.coords < 1 - 1e-9), axis=1)
    scale = np.abs(f).max()
    assert np.abs(f[interior]).max() < 1e-9 * scale


@pytest.mark.parametrize("name", list(FAMS))
def test_dense_bruteforce_assembly(port, name):
    m = meshgen.block(2, FAMS[name], seed_geo=10, seed_E=11)
    dof_eq, nf, _ = m.dof_map()
    K = np.zeros((nf, nf))
    for e, nodes in enumerate(m.blocks[0][1]):
        Ke = port.element_stiffness(FAMS[name], m.coords[nodes], m.E[e], m.nu[e])
        d = np.array([dof_eq[3 * n + c] for n in nodes for c in range(3)])
        K[np.ix_(d, d)] += Ke
    s = port.assemble_linear_static(m)
    rp, ci, v = s.csr()
    rows = np.repeat(np.arange(nf), np.diff(rp))
    assert np.array_equal(np.sort(np.nonzero(K.ravel())[0]),
                          np.sort(rows.astype(np.int64) * nf + ci))
    assert rel(v, K[rows, ci]) < 1e-14


def test_dof_map_and_loads_match_oracle(port):
    m = meshgen.cantilever()
    dof_eq, nf, pres = m.dof_map()
    s = port.assemble_linear_static(m)
    assert nf == s.n_free and np.array_equal(dof_eq, s.dof_eq)
    f = m.external_load()
    assert np.array_equal(f[dof_eq >= 0], s.rhs)  # no prescribed values -> rhs = loads
    assert abs(f.sum() + 0.04) < 1e-15


J2_ISO = J2Params(210000.0, 0.3, [0.0, 1.0], [250.0, 2250.0])


def _mu(p):
    return p.E / (2 * (1 + p.nu))


def test_j2_elastic_limit(port):
    lam = J2_ISO.E * J2_ISO.nu / ((1 + J2_ISO.nu) * (1 - 2 * J2_ISO.nu))
    mu = _mu(J2_ISO)
    De = np.zeros((6, 6))
    De[:3, :3] = lam
    De[np.arange(3), np.arange(3)] += 2 * mu
    De[np.arange(3, 6), np.arange(3, 6)] = mu
    eps = np.array([1e-5, -2e-5, 3e-6, 4e-6, -1e-6, 2e-6])
    sig, D, st = port.j2_evaluate(J2_ISO, eps, np.zeros(26))
    assert np.abs(D - De).max() < 1e-9 * mu
    assert rel(sig, De @ eps) < 1e-13
    assert np.all(st[13:] == 0)


def _vm(s):
    dev = s.copy()
    dev[:3] -= s[:3].mean()
    return np.sqrt(1.5 * (dev[:3] @ dev[:3] + 2 * dev[3:] @ dev[3:]))


def test_j2_return_on_surface_and_closed_form(port):
    rng = np.random.default_rng(12)
    mu, H = _mu(J2_ISO), 2000.0
    for _ in range(20):
        eps = rng.normal(scale=4e-3, size=6)
        st0 = np.zeros(26)
        st0[6] = rng.uniform(0, 1e-3)
        sig, D, st = port.j2_evaluate(J2_ISO, eps, st0)
        dg = st[19] - st0[6]
        if dg == 0.0:
            continue
        sy = 250.0 + H * st[19]
        assert abs(_vm(sig) - sy) < 1e-9 * sy  # lands on the hardened yield surface
        # linear isotropic hardening: dgamma = f_trial / (3 mu + H)
        lam = J2_ISO.E * J2_ISO.nu / ((1 + J2_ISO.nu) * (1 - 2 * J2_ISO.nu))
        tr = lam * eps[:3].sum()
        s_tr = np.concatenate([tr + 2 * mu * eps[:3], mu * eps[3:]])
        f_tr = _vm(s_tr) - (250.0 + H * st0[6])
        assert abs(dg - f_tr / (3 * mu + H)) < 1e-12 * dg


@pytest.mark.parametrize("hard", [_lib.HARD_ISOTROPIC, _lib.HARD_KINEMATIC, _lib.HARD_COMBINED])
def test_j2_consistent_tangent_fd(port, hard):
    p = J2Params(200000.0, 0.3, [0.0, 0.02, 0.2], [300.0, 400.0, 450.0], hard)
    rng = np.random.default_rng(13)
    hist = meshgen.seeded_history(1, 14)[:, 0]
    st0 = np.concatenate([hist, hist])
    for _ in range(6):
        eps = rng.normal(scale=3e-3, size=6)
        _, D, st = port.j2_evaluate(p, eps, st0)
        if st[19] == st0[6]:
            continue
        h = 1e-8
        Dfd = np.zeros((6, 6))
        for j in range(6):
            e = np.zeros(6)
            e[j] = h
            sp, _, _ = port.j2_evaluate(p, eps + e, st0)
            sm, _, _ = port.j2_evaluate(p, eps - e, st0)
            Dfd[:, j] = (sp - sm) / (2 * h)
        assert np.abs(D - Dfd).max() < 1e-5 * np.abs(D).max()


def test_vjp_central_differences(port):
    m, rho = meshgen.adjoint_block(3, 2, 2, (1.5, 1.0, 1.0))
    u = meshgen.smooth_field(m.coords, 7)
    w = np.random.default_rng(8).normal(size=(m.n_nodes, 3))
    dx, dE, du, r = port.vjp(m, u, w)
    assert rel(r, port.internal_force(m, u)) < 1e-13

    def phi(coords=None, E=None):
        mm = Model(m.coords if coords is None else coords, m.blocks,
                   m.E if E is None else E, m.nu)
        return w.ravel() @ port.internal_force(mm, u)
    h = 1e-6
    fd = np.zeros_like(m.coords)
    for i in range(m.n_nodes):
        for c in range(3):
            xp, xm = m.coords.copy(), m.coords.copy()
            xp[i, c] += h
            xm[i, c] -= h
            fd[i, c] = (phi(coords=xp) - phi(coords=xm)) / (2 * h)
    assert rel(dx, fd) < 1e-6
    # d phi / dE_e = phi_e / E_e (stiffness linear in E)
    fdE = np.zeros(m.n_elem)
    for e in range(m.n_elem):
        Ep, Em = m.E.copy(), m.E.copy()
        Ep[e] *= 1 + 1e-6
        Em[e] *= 1 - 1e-6
        fdE[e] = (phi(E=Ep) - phi(E=Em)) / (2e-6 * m.E[e])
    assert rel(dE, fdE) < 1e-6
    # du = K w with the assembled operator (no BCs)
    s = port.assemble_linear_static(m)
    assert rel(du, s.dense() @ w.ravel()) < 1e-12


# ---- field recovery (recover_fields / recover_gauss_stress) ------------------------

def _recover_fixture_models():
    for name, t in FAMS.items():
        yield name, meshgen.block(2, t, seed_geo=21, seed_E=22)


def test_golden_recover(port):
    """recover_fields / recover_gauss_stress against the reference core's outputs."""
    from oracle.oracle import aos_state
    g = gold("recover")
    for name, m in _recover_fixture_models():
        o = port.recover_fields(m, g[name + "_u"], gauss_stress=True)
        for k in ("strain", "stress", "reaction", "gauss_stress"):
            assert np.array_equal(o[k], g[name + "_" + k]), (name, k)
    m = meshgen.cantilever()
    o = port.recover_fields(m, g["cantilever_u"])
    for k in ("strain", "stress", "reaction"):
        assert np.array_equal(o[k], g["cantilever_" + k]), k
    m = meshgen.j2_block(3, 2, 2, (1.0, 0.8, 0.6), family=_lib.C3D20, elastic_every=5,
                         jitter_frac=0.1)
    o = port.recover_fields(m, g["j2_u"], aos_state(g["j2_history"]))
    for k in ("strain", "stress", "reaction"):
        assert np.array_equal(o[k], g["j2_" + k]), k


@pytest.mark.parametrize("name", list(FAMS))
def test_reference_bitwise_recover(port, ref, name):
    from oracle.oracle import aos_state
    m = meshgen.block(3, FAMS[name], seed_geo=111, seed_E=112, ny=2, nz=2)
    u = meshgen.smooth_field(m.coords, 113)
    a, b = port.recover_fields(m, u, gauss_stress=True), ref.recover_fields(m, u, gauss_stress=True)
    for k in ("strain", "stress", "reaction", "gauss_stress"):
        assert np.array_equal(a[k], b[k]), k
    m = meshgen.j2_block(3, 2, 2, (1.0, 0.8, 0.6), family=FAMS[name], elastic_every=4,
                         jitter_frac=0.1)
    h = aos_state(meshgen.seeded_history(m.n_points, 114))
    u = meshgen.smooth_field(m.coords, 115, 1.5e-4)
    a, b = port.recover_fields(m, u, h), ref.recover_fields(m, u, h)
    for k in ("strain", "stress", "reaction"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("name", list(FAMS))
def test_recover_linear_field_closed_form(port, name):
    """u = A x + b: every Gauss point, element and node carries the strain of A
    (engineering shear) and D of it; the internal forces sum to zero."""
    jittered = meshgen.block(3, FAMS[name], seed_geo=12, ny=2, nz=2,
                             jitter_frac=0.0 if name == "C3D20R" else 0.1)
    E, nu = 2.0e5, 0.3
    m = Model(jittered.coords, jittered.blocks, E, nu)
    rng = np.random.default_rng(13)
    A = rng.normal(size=(3, 3)) * 1e-3
    u = m.coords @ A.T + rng.normal(size=3) * 1e-3
    eps = np.array([A[0, 0], A[1, 1], A[2, 2], A[0, 1] + A[1, 0], A[0, 2] + A[2, 0],
                    A[1, 2] + A[2, 1]])
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    sig = np.concatenate([lam * eps[:3].sum() + 2 * mu * eps[:3], mu * eps[3:]])
    o = port.recover_fields(m, u, gauss_stress=True)
    assert np.abs(o["strain"] - eps).max() <= 1e-12 * np.abs(eps).max()
    assert np.abs(o["stress"] - sig).max() <= 1e-12 * np.abs(sig).max()
    assert np.abs(o["gauss_stress"] - sig).max() <= 1e-12 * np.abs(sig).max()
    f = o["f_int"]
    assert np.abs(f.sum(axis=0)).max() <= 1e-12 * np.abs(f).max()
    assert np.array_equal(o["reaction"], f)  # no loads


def test_recover_node_average_weights(port):
    """Two elements of different stiffness under the same uniform strain: a shared
    node's stress is the plain mean of the two elements' stresses (counted once per
    element), an unshared node's is its own element's."""
    c, conn = meshgen.hex_grid(2, 1, 1, (2.0, 1.0, 1.0), _lib.C3D8)
    E = np.array([1.0e5, 3.0e5])
    m = Model(c, [(_lib.C3D8, conn)], E, 0.25)
    u = np.zeros_like(c)
    u[:, 0] = 1e-3 * c[:, 0]
    o = port.recover_fields(m, u)
    s_e = [port.recover_fields(Model(c, [(_lib.C3D8, conn[i:i + 1])], E[i], 0.25), u)
           ["stress"][conn[i][0]] for i in range(2)]
    shared = np.intersect1d(conn[0], conn[1])
    only0 = np.setdiff1d(conn[0], conn[1])
    assert np.allclose(o["stress"][shared], 0.5 * (s_e[0] + s_e[1]), rtol=1e-14, atol=0)
    assert np.allclose(o["stress"][only0], s_e[0], rtol=1e-14, atol=0)
    assert np.all(o["stress"][shared, 0] > s_e[0][0]) and np.all(o["stress"][shared, 0] < s_e[1][0])


# ---- mass and geometric stiffness (assemble_mass, assemble_geometric_stiffness) -------

def _same_system(a, exp_prefix, g):
    rp, ci, v = a.csr()
    assert np.array_equal(rp, g[exp_prefix + "row_ptr"])
    assert np.array_equal(ci, g[exp_prefix + "col_idx"])
    assert np.array_equal(v, g[exp_prefix + "vals"])
    assert np.array_equal(a.rhs, g[exp_prefix + "rhs"])


def test_golden_mass_geometric(port):
    g = gold("mass_geometric")
    for name, t in FAMS.items():
        b = meshgen.block(2, t, seed_geo=21, seed_E=22)
        m = Model(b.coords, b.blocks, b.E, b.nu, rho=g[name + "_rho"])
        _same_system(port.assemble_mass(m), name + "_mass_", g)
        _same_system(port.assemble_mass(m, True), name + "_lumped_", g)
        _same_system(port.assemble_geometric_stiffness(m, g[name + "_gp_stress"]), name + "_geo_", g)


@pytest.mark.parametrize("name", list(FAMS))
def test_reference_bitwise_mass_geometric(port, ref, name):
    b = meshgen.cantilever(4, 2, 2) if name == "C3D8" else meshgen.block(3, FAMS[name], seed_geo=131)
    rng = np.random.default_rng(132)
    m = Model(b.coords, b.blocks, b.E, b.nu, b.spc_node, b.spc_comp, b.spc_value,
              rho=rng.uniform(1e-9, 8e-9, b.n_elem))
    gs = rng.normal(scale=50.0, size=(m.n_points, 6))
    for a, r in ((port.assemble_mass(m), ref.assemble_mass(m)),
                 (port.assemble_mass(m, True), ref.assemble_mass(m, True)),
                 (port.assemble_geometric_stiffness(m, gs), ref.assemble_geometric_stiffness(m, gs))):
        for x, y in zip(a.csr(), r.csr()):
            assert np.array_equal(x, y)
        assert np.array_equal(a.rhs, r.rhs) and not a.rhs.any()


@pytest.mark.parametrize("name", list(FAMS))
def test_mass_closed_forms(port, name):
    """Total mass rho V on each direction (partition of unity), only same-direction
    couplings, and the lumped diagonal equal to the consistent row sums. (The reduced
    C3D20R rule integrates det J exactly only on undistorted elements.)"""
    b = meshgen.block(3, FAMS[name], seed_geo=14, ny=2, nz=2,
                      jitter_frac=0.0 if name == "C3D20R" else 0.1)  # boundary stays put
    L = b.coords.max(axis=0) - b.coords.min(axis=0)
    rho = 7.85e-9
    m = Model(b.coords, b.blocks, b.E, b.nu, rho=rho)
    M = port.assemble_mass(m).dense()
    V = float(np.prod(L))
    for d in range(3):
        assert abs(M[d::3, d::3].sum() - rho * V) <= 1e-13 * rho * V
        for d2 in range(3):
            if d2 != d:
                assert not M[d::3, d2::3].any()
    Ml = port.assemble_mass(m, True).dense()
    assert np.array_equal(Ml, np.diag(np.diag(Ml)))
    assert np.allclose(np.diag(Ml), M.sum(axis=1), rtol=1e-13, atol=0)


@pytest.mark.parametrize("name", list(FAMS))
def test_geometric_closed_forms(port, name):
    """K_geo annihilates rigid translations, is symmetric and linear in the stress,
    and couples only equal directions."""
    b = meshgen.block(3, FAMS[name], seed_geo=15, ny=2, nz=2)
    m = Model(b.coords, b.blocks, b.E, b.nu)
    rng = np.random.default_rng(16)
    s1, s2 = rng.normal(size=(m.n_points, 6)), rng.normal(size=(m.n_points, 6))
    K1 = port.assemble_geometric_stiffness(m, s1).dense()
    K2 = port.assemble_geometric_stiffness(m, s2).dense()
    K12 = port.assemble_geometric_stiffness(m, s1 + 2.0 * s2).dense()
    scale = np.abs(K12).max()
    assert np.abs(K12 - (K1 + 2.0 * K2)).max() <= 1e-12 * scale
    assert np.array_equal(K1, K1.T)
    for d in range(3):
        t = np.zeros(K1.shape[0])
        t[d::3] = 1.0
        assert np.abs(K1 @ t).max() <= 1e-12 * np.abs(K1).max()
    assert not K1[0::3, 1::3].any()


def test_geometric_unit_cube_entry(port):
    """One undistorted unit C3D8 cube under s_xx: K_geo[0, 0] = s_xx * integral of
    (dN_1/dx)^2 = s_xx (1/3)(1/3)."""
    c, conn = meshgen.hex_grid(1, 1, 1, (1.0, 1.0, 1.0), _lib.C3D8)
    m = Model(c, [(_lib.C3D8, conn)], 2.0e5, 0.3)
    s = np.zeros((8, 6))
    s[:, 0] = 90.0
    K = port.assemble_geometric_stiffness(m, s).dense()
    assert abs(K[0, 0] - 10.0) <= 1e-12
