"""Edge cases of the path: nodes that no element uses (free DOFs with empty rows, one of
them loaded), a model whose every DOF is prescribed (no free equation, no nonzero), and a
single element of each family.

The oracle is checked against the reference core bit for bit on these models (CPU, needs
oracle/_ref); the device path is checked against the oracle (GPU):
  CSR row_ptr / col_idx   bit-exact
  K, mass values, rhs     rel-L2 <= 1e-12
  f_int, recovered fields rel-L2 <= 1e-12; at nodes of no element f_int, strain and
                          stress exactly zero and the reaction minus the node's load
"""
import numpy as np
import pytest

REL = 1e-12


def rel_l2(a, b):
    a, b = np.ravel(a), np.ravel(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _families():
    from calculixpp_amd import _lib
    return [_lib.C3D8, _lib.C3D20, _lib.C3D20R, _lib.C3D6]


def unused_nodes_model(fam, seed):
    """A 2^3 block with one unused node before its nodes and two after; the x = 0 face
    clamped, a z load on x = 1 and loads on the first and the last (unused) node."""
    from calculixpp_amd import meshgen
    from calculixpp_amd.model import Model
    b = meshgen.block(2, fam, seed_geo=seed, seed_E=seed + 1)
    coords = np.vstack([[[5.0, 5.0, 5.0]], b.coords, [[6.0, 5.0, 5.0], [-3.0, 0.5, 0.25]]])
    blocks = [(t, c + 1) for t, c in b.blocks]
    x0 = np.nonzero(coords[:, 0] == 0.0)[0]
    x1 = np.nonzero(coords[:, 0] == 1.0)[0]
    last = coords.shape[0] - 1
    rng = np.random.default_rng(seed + 2)
    return Model(coords, blocks, b.E, b.nu,
                 spc_node=np.repeat(x0, 3), spc_comp=np.tile([1, 2, 3], len(x0)),
                 spc_value=np.zeros(3 * len(x0)),
                 cload_node=np.append(x1, [0, last]),
                 cload_comp=np.append(np.full(len(x1), 3), [1, 2]),
                 cload_value=np.append(np.full(len(x1), -1.0), [2.0, -0.5]),
                 rho=rng.uniform(1e-9, 8e-9, b.n_elem))


def clamped_model(fam, seed):
    """Every DOF prescribed (nonzero values) and one load: no free equation."""
    from calculixpp_amd import meshgen
    from calculixpp_amd.model import Model
    b = meshgen.block(2, fam, seed_geo=seed, seed_E=seed + 1, ny=1, nz=1)
    n = b.n_nodes
    return Model(b.coords, b.blocks, b.E, b.nu,
                 spc_node=np.repeat(np.arange(n), 3), spc_comp=np.tile([1, 2, 3], n),
                 spc_value=np.linspace(0.0, 1e-3, 3 * n),
                 cload_node=np.array([0]), cload_comp=np.array([1]), cload_value=np.array([1.0]),
                 rho=np.full(b.n_elem, 7.85e-9))


def single_element_model(fam, seed):
    """One jittered element, its x = 0 face clamped, a z load on x = 1."""
    from calculixpp_amd import meshgen
    from calculixpp_amd.model import Model
    b = meshgen.block(1, fam, seed_geo=seed, seed_E=seed + 1, jitter_frac=0.0)
    rng = np.random.default_rng(seed + 2)
    coords = b.coords + rng.uniform(-0.05, 0.05, b.coords.shape)
    x0 = np.nonzero(b.coords[:, 0] == 0.0)[0]
    x1 = np.nonzero(b.coords[:, 0] == 1.0)[0]
    return Model(coords, b.blocks, b.E, b.nu,
                 spc_node=np.repeat(x0, 3), spc_comp=np.tile([1, 2, 3], len(x0)),
                 spc_value=np.zeros(3 * len(x0)),
                 cload_node=x1, cload_comp=np.full(len(x1), 3), cload_value=np.full(len(x1), -1.0),
                 rho=np.full(b.n_elem, 7.85e-9))


MODELS = {"unused_nodes": unused_nodes_model, "clamped": clamped_model,
          "single_element": single_element_model}


def _unused(model):
    used = np.zeros(model.n_nodes, bool)
    for _, c in model.blocks:
        used[np.unique(c)] = True
    return ~used


@pytest.mark.parametrize("kind", sorted(MODELS))
@pytest.mark.parametrize("fam_i", range(4))
def test_oracle_matches_reference_core(port, ref, kind, fam_i):
    """The restatement reproduces the reference core bit for bit on the edge models:
    K, the mass matrix, f_int and the recovered fields."""
    from calculixpp_amd import meshgen
    m = MODELS[kind](_families()[fam_i], 81 + fam_i)
    u = meshgen.smooth_field(m.coords, 83)
    for a, b in ((port.assemble_linear_static(m), ref.assemble_linear_static(m)),
                 (port.assemble_mass(m), ref.assemble_mass(m)),
                 (port.assemble_mass(m, True), ref.assemble_mass(m, True))):
        assert a.n_free == b.n_free
        for x, y in zip(a.csr(), b.csr()):
            assert np.array_equal(x, y)
        assert np.array_equal(a.rhs, b.rhs)
    assert np.array_equal(port.internal_force(m, u), ref.internal_force(m, u))
    rp, rr = port.recover_fields(m, u), ref.recover_fields(m, u)
    for k in ("strain", "stress", "reaction"):
        assert np.array_equal(rp[k], rr[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("kind", sorted(MODELS))
@pytest.mark.parametrize("fam_i", range(4))
def test_device_edge_models(port, kind, fam_i):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from calculixpp_amd import fem, meshgen, recover
    m = MODELS[kind](_families()[fam_i], 81 + fam_i)
    u = meshgen.smooth_field(m.coords, 83)
    unused = _unused(m)
    for got, exp in ((fem.assemble_linear_static(m), port.assemble_linear_static(m)),
                     (fem.assemble_mass(m), port.assemble_mass(m)),
                     (fem.assemble_mass(m, True), port.assemble_mass(m, True))):
        rp, ci, v = exp.csr()
        assert got.n_free == exp.n_free
        assert np.array_equal(got.row_ptr.cpu().numpy(), rp)
        assert np.array_equal(got.col_idx.cpu().numpy(), ci)
        assert rel_l2(got.vals.cpu().numpy()[: len(v)], v) <= REL
        assert rel_l2(got.rhs.cpu().numpy(), exp.rhs) <= REL
    f = fem.internal_force(m, u).cpu().numpy().reshape(-1, 3)
    assert rel_l2(f, port.internal_force(m, u)) <= REL
    assert not f[unused].any()
    r = recover.FieldRecovery(m)
    o = {k: v.cpu().numpy() for k, v in r.run(torch.from_numpy(u)).items()}
    exp = port.recover_fields(m, u)
    for k in ("strain", "stress", "reaction"):
        assert rel_l2(o[k], exp[k]) <= REL, k
    for k in ("strain", "stress"):
        assert not o[k].reshape(m.n_nodes, -1)[unused].any(), k
    # reaction = f_int - f_ext: at a node of no element, minus its load
    f_ext = np.zeros((m.n_nodes, 3))
    np.add.at(f_ext, (m.cload_node, m.cload_comp - 1), m.cload_value)
    assert np.array_equal(o["reaction"].reshape(-1, 3)[unused], -f_ext[unused])


def _collapsed_hex_model():
    """A 2x1x1 C3D8 block whose second element has edge 2-3 collapsed (node 3 := node 2),
    a wedge-shaped hex with a positive Jacobian at every Gauss point. Built valid, then
    collapsed in place: the array model itself refuses such elements."""
    from calculixpp_amd import _lib, meshgen
    from calculixpp_amd.model import Model
    coords, conn = meshgen.hex_grid(2, 1, 1, (1.0, 0.5, 0.5), _lib.C3D8)
    m = Model(coords, [(_lib.C3D8, conn)], 210000.0, 0.3)
    m.blocks[0][1][1, 3] = m.blocks[0][1][1, 2]
    return m


def test_model_refuses_collapsed_element():
    from calculixpp_amd import _lib, meshgen
    from calculixpp_amd.model import Model
    coords, conn = meshgen.hex_grid(2, 1, 1, (1.0, 0.5, 0.5), _lib.C3D8)
    conn[1, 7] = conn[1, 6]
    with pytest.raises(ValueError, match="collapsed element"):
        Model(coords, [(_lib.C3D8, conn)], 210000.0, 0.3)


def test_reference_assembles_the_collapsed_element(ref):
    """The reference accepts the collapsed hex (the mesh the device path refuses)."""
    m = _collapsed_hex_model()
    s = ref.assemble_linear_static(m)
    assert s.n_free == 3 * m.n_nodes and np.isfinite(s.vals).all()


@pytest.mark.gpu
def test_device_refuses_collapsed_element():
    """cxg_mesh_create refuses an element that lists a node twice (the row gather and the
    node sums hold one incidence per node and element), and so does the C++ drop-in on the
    reference's own Model, with the same message: no wrong numbers come back."""
    import ctypes as C
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from calculixpp_amd import _lib, fem
    m = _collapsed_hex_model()
    ctx = fem.context()
    conn = torch.from_numpy(m.blocks[0][1]).to(ctx.device)
    blocks = (_lib.Block * 1)()
    blocks[0].type, blocks[0].n_elem = _lib.C3D8, conn.shape[0]
    blocks[0].elem_offset, blocks[0].gp_offset, blocks[0].conn = 0, 0, conn.data_ptr()
    h = C.c_void_p()
    st = ctx.lib.cxg_mesh_create(ctx.h, m.n_nodes, 1, blocks, C.byref(h))
    assert st == _lib.ERR_UNSUPPORTED and not h.value
    assert b"collapsed element" in ctx.lib.cxg_last_error(ctx.h)
    from oracle.oracle import Oracle, OracleError, dropin_available
    if dropin_available():
        with pytest.raises(OracleError, match="collapsed element"):
            Oracle("dropin").assemble_linear_static(m)
