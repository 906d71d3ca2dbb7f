"""Mass and geometric stiffness on the device (cxg_assemble_mass, cxg_assemble_geometric
via fem.assemble_mass / fem.assemble_geometric_stiffness) against the oracle, which is
pinned bit for bit to the reference core (tests/test_oracle.py):
  CSR row_ptr / col_idx   bit-exact (the reference's key set: only same-direction
                          couplings, entries with every contribution 0.0 left out)
  values                  rel-L2 <= 1e-12
  rhs                     zero (the eliminated-column side effect is discarded)
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL = 1e-12


def rel_l2(a, b):
    a, b = np.ravel(a), np.ravel(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from calculixpp_amd import fem
    return fem


def _families():
    from calculixpp_amd import _lib
    return [_lib.C3D8, _lib.C3D20, _lib.C3D20R, _lib.C3D6]


def same(sys_g, sys_e):
    rp, ci, v = sys_e.csr()
    assert sys_g.n_free == sys_e.n_free
    assert np.array_equal(sys_g.row_ptr.cpu().numpy(), rp)
    assert np.array_equal(sys_g.col_idx.cpu().numpy(), ci)
    assert rel_l2(sys_g.vals.cpu().numpy(), v) <= REL, rel_l2(sys_g.vals.cpu().numpy(), v)
    assert not sys_g.rhs.cpu().numpy().any()


def _model(fam, seed, spc=False):
    from calculixpp_amd import meshgen
    from calculixpp_amd.model import Model
    b = meshgen.block(4, fam, seed_geo=seed, seed_E=seed + 1, ny=3, nz=2)
    kw = {}
    if spc:
        x0 = np.nonzero(b.coords[:, 0] < 1e-12)[0]
        kw = dict(spc_node=np.repeat(x0, 2), spc_comp=np.tile([1, 3], len(x0)),
                  spc_value=np.tile([0.0, 1e-3], len(x0)))
    rng = np.random.default_rng(seed + 2)
    return Model(b.coords, b.blocks, b.E, b.nu, rho=rng.uniform(1e-9, 8e-9, b.n_elem), **kw)


@pytest.mark.parametrize("lumped", [False, True])
@pytest.mark.parametrize("fam_i", range(4))
def test_mass(gpu, port, fam_i, lumped):
    m = _model(_families()[fam_i], 141, spc=True)
    same(gpu.assemble_mass(m, lumped), port.assemble_mass(m, lumped))


@pytest.mark.parametrize("fam_i", range(4))
def test_geometric_stiffness(gpu, port, fam_i):
    m = _model(_families()[fam_i], 151, spc=True)
    gs = np.random.default_rng(152).normal(scale=80.0, size=(m.n_points, 6))
    same(gpu.assemble_geometric_stiffness(m, gs), port.assemble_geometric_stiffness(m, gs))


def test_geometric_from_recovered_stress_and_removed_elements(gpu, port):
    """The buckling prestress chain: recover_gauss_stress of a displacement feeds
    assemble_geometric_stiffness, on a model with removed elements (their points are
    empty and they add no mass or K_geo)."""
    from calculixpp_amd import meshgen, recover
    m = meshgen.cantilever()
    m.rho = np.full(m.n_elem, 7.85e-9)
    m.active = np.ones(m.n_elem, bool)
    m.active[1::4] = False
    u = meshgen.smooth_field(m.coords, 161)
    gs = recover.recover_gauss_stress(m, u)
    exp_gs = port.recover_fields(m, u, gauss_stress=True)["gauss_stress"]
    same(gpu.assemble_geometric_stiffness(m, gs), port.assemble_geometric_stiffness(m, exp_gs))
    same(gpu.assemble_mass(m), port.assemble_mass(m))


def test_zero_density_and_zero_stress_give_empty_systems(gpu):
    from calculixpp_amd import _lib, meshgen
    m = meshgen.block(2, _lib.C3D20, seed_geo=171)
    s = gpu.assemble_mass(m)  # no *DENSITY
    assert s.vals.numel() == 0 and s.row_ptr.cpu().numpy().max() == 0
    s = gpu.assemble_geometric_stiffness(m, np.zeros((m.n_points, 6)))
    assert s.vals.numel() == 0


@pytest.mark.parametrize("kind", ["mass", "lumped", "geometric"])
@pytest.mark.parametrize("mesh", ["c3d20", "mixed"])
def test_same_direction_rows_equal_structural_then_compacted(gpu, port, mesh, kind):
    """The same-direction gather (rows written straight onto the (a, i)-(b, i) pattern)
    gives the key set of the structural gather + compaction bit for bit: pattern, values
    and liveness. Its pattern is the structural one with each row's columns of the other
    two directions dropped. gather_lds_cols 8 (rows longer than the LDS) takes the
    structural path in fem.assemble_mass, with the same result."""
    import torch
    from calculixpp_amd import _lib, meshgen
    from calculixpp_amd.model import Model
    if mesh == "mixed":
        coords, blocks = meshgen.mixed_box(700, 181)
        n_el = sum(len(c) for _, c in blocks)
        rng = np.random.default_rng(182)
        x0 = np.nonzero(coords[:, 0] < 1e-12)[0]
        m = Model(coords, blocks, rng.uniform(1e5, 2e5, n_el), 0.3, rho=rng.uniform(1e-9, 8e-9, n_el),
                  spc_node=x0, spc_comp=np.full(len(x0), 2), spc_value=np.zeros(len(x0)))
    else:
        m = _model(_lib.C3D20, 183, spc=True)
    gs = torch.from_numpy(np.random.default_rng(184).normal(scale=80.0, size=(m.n_points, 6)))
    ctx = gpu.context()
    a = gpu.Assembler(m, ctx)
    rp_s, ci_s = a.pattern.row_ptr.cpu().numpy(), a.pattern.col_idx.cpu().numpy()
    rp_d, ci_d = (t.cpu().numpy() for t in a.same_direction())
    comp = np.empty(a.n_free, np.int64)
    eq = np.asarray(a.dof_eq)
    comp[eq[eq >= 0]] = np.nonzero(eq >= 0)[0] % 3
    rows = [ci_s[rp_s[r]:rp_s[r + 1]] for r in range(a.n_free)]
    keep = [c[comp[c] == comp[r]] for r, c in enumerate(rows)]
    assert np.array_equal(rp_d, np.concatenate([[0], np.cumsum([len(k) for k in keep])]))
    assert np.array_equal(ci_d, np.concatenate(keep) if keep else ci_d[:0])
    dev = ctx.device
    nnz_d = max(len(ci_d), 1)
    v_d = torch.empty(nnz_d, dtype=torch.float64, device=dev)
    l_d = torch.empty(nnz_d, dtype=torch.uint8, device=dev)
    v_s, l_s = a.new_values(), a.new_live()
    if kind == "geometric":
        a.assemble_geometric_sd(v_d, l_d, gs)
        a.assemble_geometric(v_s, gs, live=l_s)
    else:
        a.assemble_mass_sd(v_d, l_d, kind == "lumped")
        a.assemble_mass(v_s, kind == "lumped", live=l_s)
    rd, cd, vd = (t.cpu().numpy() for t in a.sd_system(v_d, l_d))
    rs_, cs_, vs_ = (t.cpu().numpy() for t in a.compact(v_s, l_s))
    assert np.array_equal(rd, rs_) and np.array_equal(cd, cs_) and np.array_equal(vd, vs_)
    # the structural values off the same-direction pattern are exact zeros, never live
    ls = l_s[: a.pattern.nnz].cpu().numpy()
    assert ls.sum() == l_d[: len(ci_d)].cpu().numpy().sum()
    try:
        ctx.set_option("gather_lds_cols", 8)
        b = gpu.Assembler(m, ctx)
        assert not b.sd_fits()
        fallback = (gpu.assemble_geometric_stiffness(m, gs.numpy()) if kind == "geometric"
                    else gpu.assemble_mass(m, kind == "lumped"))
    finally:
        ctx.set_option("gather_lds_cols", 256)
    assert np.array_equal(fallback.row_ptr.cpu().numpy(), rd)
    assert np.array_equal(fallback.col_idx.cpu().numpy(), cd)
    assert np.array_equal(fallback.vals.cpu().numpy(), vd)
    exp = (port.assemble_geometric_stiffness(m, gs.numpy()) if kind == "geometric"
           else port.assemble_mass(m, kind == "lumped"))
    same(gpu.assemble_geometric_stiffness(m, gs.numpy()) if kind == "geometric"
         else gpu.assemble_mass(m, kind == "lumped"), exp)
