"""J2 hardening variants on the device against the oracle and the reference core.

The radial return of src/fem/plasticity.cpp:90-236 carries a kinematic modulus H_kin
(derived from the first table slope for KINEMATIC and COMBINED hardening when the deck
gives none, with the isotropic curve flattened for pure KINEMATIC, plasticity.cpp:51-62)
and a piecewise-linear yield table with flat extrapolation past its last point
(plasticity.cpp:67-88). Each case here runs a seeded committed history through one
material-tangent assembly:
  - the Python path (libcxg.so through calculixpp_amd.fem) against the oracle: CSR
    pattern bit-exact, K_t and f_int rel-L2 <= 1e-12, trial history rel-L2 <= 1e-10;
  - the C++ drop-in cxpp::gpu::assemble_material_tangent against the reference's own
    fem::assemble_material_tangent on the same cxpp::Model (same bars);
and checks that the case really exercises what it names (yielding points, points past
the last table point, points in more than one table segment, a nonzero back-stress
update).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL_K = 1e-12
REL_STATE = 1e-10

# name -> (hardening, eqplastic, yield, kinematic_modulus)
TABLES = {
    # two points: H = 2000 (the config-3 table) as kinematic / combined hardening
    "kin2": ("kin", [0.0, 1.0], [250.0, 2250.0], 0.0),
    "comb2": ("comb", [0.0, 1.0], [250.0, 2250.0], 0.0),
    # four points, the last at 1.6e-3 (the seeded history reaches ep = 2e-3, so about a
    # fifth of the points start past it and take the flat extrapolation)
    "iso4": ("iso", [0.0, 4e-4, 1e-3, 1.6e-3], [250.0, 290.0, 320.0, 335.0], 0.0),
    "kin4": ("kin", [0.0, 4e-4, 1e-3, 1.6e-3], [250.0, 290.0, 320.0, 335.0], 0.0),
    "comb4": ("comb", [0.0, 4e-4, 1e-3, 1.6e-3], [250.0, 290.0, 320.0, 335.0], 0.0),
    # combined hardening with the kinematic modulus given explicitly (no table slope)
    "comb4_hk": ("comb", [0.0, 4e-4, 1e-3, 1.6e-3], [250.0, 290.0, 320.0, 335.0], 5000.0),
}


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from calculixpp_amd import fem
    return fem


def _model(family, table, elastic_every=5):
    from calculixpp_amd import _lib, meshgen
    from calculixpp_amd.model import J2Params
    hard, ep, sy, hk = TABLES[table]
    kind = {"iso": _lib.HARD_ISOTROPIC, "kin": _lib.HARD_KINEMATIC,
            "comb": _lib.HARD_COMBINED}[hard]
    m = meshgen.j2_block(4, 3, 3, (1.0, 0.8, 0.6), family=family, elastic_every=elastic_every,
                         jitter_frac=0.1)
    m.j2 = [J2Params(210000.0, 0.3, list(ep), list(sy), kind, hk)]
    return m


def _exercised(table, st):
    """The trial state exercises the table: yielding points, and for the four-point
    tables points past the last table point and points in the interior segments; for
    kinematic and combined hardening a back-stress update."""
    hard, ep, _, _ = TABLES[table]
    yielded = st[:, 19] > st[:, 6]
    assert 0.05 < yielded.mean() < 1.0, yielded.mean()
    if len(ep) > 2:
        assert (yielded & (st[:, 6] >= ep[-1])).sum() > 0  # flat extrapolation
        inner = yielded & (st[:, 19] > ep[1]) & (st[:, 19] < ep[-1])
        assert inner.sum() > 0
    dal = np.abs(st[:, 20:26] - st[:, 7:13]).max(axis=1)
    if hard == "iso":
        assert dal.max() == 0.0
    else:
        assert (dal[yielded] > 0).mean() > 0.99


@pytest.mark.parametrize("table", sorted(TABLES))
@pytest.mark.parametrize("fam_i", range(4))
def test_hardening_python_path(gpu, port, fam_i, table):
    from calculixpp_amd import _lib, meshgen
    from oracle.oracle import aos_state
    fam = [_lib.C3D8, _lib.C3D20, _lib.C3D20R, _lib.C3D6][fam_i]
    m = _model(fam, table)
    hist = meshgen.seeded_history(m.n_points, 3)
    u = meshgen.smooth_field(m.coords, 7, 1.5e-4)
    sys_e, f_e, st_e = port.assemble_material_tangent(m, u, aos_state(hist))
    _exercised(table, st_e)
    mp = gpu.MaterialPoints(m.n_points, committed=hist)
    sys_g, f_g = gpu.assemble_material_tangent(m, u, mp)
    rp, ci, vals = sys_e.csr()
    assert np.array_equal(sys_g.row_ptr.cpu().numpy(), rp)
    assert np.array_equal(sys_g.col_idx.cpu().numpy(), ci)
    assert rel_l2(sys_g.vals.cpu().numpy(), vals) <= REL_K
    assert rel_l2(f_g.cpu().numpy(), f_e) <= REL_K
    trial = mp.trial.cpu().numpy()
    for c in range(13):
        assert rel_l2(trial[c], st_e[:, 13 + c]) <= REL_STATE, c


@pytest.mark.parametrize("variant", [5, 1])
@pytest.mark.parametrize("table", ["kin4", "comb4"])
def test_hardening_other_kernels(gpu, port, table, variant):
    """The persistent work-queue kernel (5) and the colour scatter (1) run the same
    return map."""
    from calculixpp_amd import _lib, meshgen
    from oracle.oracle import aos_state
    m = _model(_lib.C3D8, table)
    hist = meshgen.seeded_history(m.n_points, 3)
    u = meshgen.smooth_field(m.coords, 7, 1.5e-4)
    sys_e, f_e, st_e = port.assemble_material_tangent(m, u, aos_state(hist))
    ctx = gpu.context()
    ctx.set_option("asm_variant", variant)
    try:
        mp = gpu.MaterialPoints(m.n_points, committed=hist)
        sys_g, f_g = gpu.assemble_material_tangent(m, u, mp)
    finally:
        ctx.set_option("asm_variant", 7)
    assert rel_l2(sys_g.vals.cpu().numpy(), sys_e.csr()[2]) <= REL_K
    assert rel_l2(f_g.cpu().numpy(), f_e) <= REL_K
    trial = mp.trial.cpu().numpy()
    for c in range(13):
        assert rel_l2(trial[c], st_e[:, 13 + c]) <= REL_STATE, c


@pytest.mark.parametrize("table", ["kin2", "comb4", "kin4", "comb4_hk"])
def test_hardening_recovery_reads_committed_history(gpu, port, table):
    """recover_fields in its material-point form returns from the committed history of
    kinematic / combined materials (stress.cpp:275-353)."""
    import torch
    from calculixpp_amd import _lib, meshgen, recover
    from oracle.oracle import aos_state
    m = _model(_lib.C3D20, table)
    hist = meshgen.seeded_history(m.n_points, 3)
    u = meshgen.smooth_field(m.coords, 7, 1.5e-4)
    exp = port.recover_fields(m, u, aos_state(hist))
    r = recover.FieldRecovery(m)
    mp = gpu.MaterialPoints(m.n_points, r.ctx.device, committed=hist)
    got = r.run(torch.from_numpy(u).to(r.ctx.device), mp)
    for k in ("stress", "strain", "reaction"):
        assert rel_l2(got[k].cpu().numpy(), exp[k]) <= REL_K, k


@pytest.fixture(scope="module")
def pair():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from oracle.oracle import Oracle, dropin_available, ref_available
    if not (ref_available() and dropin_available()):
        pytest.skip("oracle/_ref reference and drop-in builds not present")
    return Oracle("reference"), Oracle("dropin")


@pytest.mark.parametrize("table", sorted(TABLES))
@pytest.mark.parametrize("fam_i", [0, 1])
def test_hardening_dropin_against_reference(pair, fam_i, table):
    """cxpp::gpu::assemble_material_tangent against the reference core on the same
    cxpp::Model with *PLASTIC, HARDENING=KINEMATIC / COMBINED and multi-point tables."""
    from calculixpp_amd import _lib, meshgen
    from oracle.oracle import aos_state
    ref, gpu_ = pair
    fam = [_lib.C3D8, _lib.C3D20][fam_i]
    m = _model(fam, table, elastic_every=4)
    st = aos_state(meshgen.seeded_history(m.n_points, 3))
    u = meshgen.smooth_field(m.coords, 7, 1.5e-4)
    s_e, f_e, st_e = ref.assemble_material_tangent(m, u, st)
    _exercised(table, st_e)
    s_g, f_g, st_g = gpu_.assemble_material_tangent(m, u, st)
    assert s_g.n_free == s_e.n_free and np.array_equal(s_g.dof_eq, s_e.dof_eq)
    rg, cg, vg = s_g.csr()
    re_, ce, ve = s_e.csr()
    assert np.array_equal(rg, re_) and np.array_equal(cg, ce)
    assert rel_l2(vg, ve) <= REL_K
    assert rel_l2(f_g, f_e) <= REL_K
    assert np.array_equal(st_g[:, :13], st_e[:, :13])
    for c in range(13, 26):
        assert rel_l2(st_g[:, c], st_e[:, c]) <= REL_STATE, c


def _random_points(n, seed):
    """Strains around the yield strain of E = 210000, sigma_y0 = 250 and a seeded history
    (the config-3 history generator)."""
    from calculixpp_amd import meshgen
    rng = np.random.default_rng(seed)
    eps = rng.normal(scale=1.2e-3, size=(n, 6))
    eps[:, 3:] *= 1.5
    return eps, meshgen.seeded_history(n, seed + 1)


def test_material_point_evaluation(gpu, port):
    """cxg_material_evaluate (MaterialModel::evaluate at independent points) against the
    oracle's J2PlasticMaterial::evaluate point by point, for every table, and the elastic
    law at points without a material (trial history left as committed)."""
    from calculixpp_amd import _lib
    from calculixpp_amd.model import J2Params
    names = sorted(TABLES)
    kinds = {"iso": _lib.HARD_ISOTROPIC, "kin": _lib.HARD_KINEMATIC, "comb": _lib.HARD_COMBINED}
    mats = [J2Params(210000.0, 0.3, list(TABLES[t][1]), list(TABLES[t][2]), kinds[TABLES[t][0]],
                     TABLES[t][3]) for t in names]
    n = 600
    eps, hist = _random_points(n, 17)
    mat_id = np.arange(n, dtype=np.int32) % (len(mats) + 1) - 1  # -1: elastic
    E = np.full(n, 190000.0)
    nu = np.full(n, 0.28)
    sig, D, trial = gpu.evaluate_material_points(eps, mats, mat_id, E, nu, hist)
    es, eD, et = np.zeros((n, 6)), np.zeros((n, 6, 6)), np.zeros((13, n))
    yielded = 0
    for p in range(n):
        st = np.zeros(26)
        st[:13] = hist[:, p]
        st[13:] = hist[:, p]
        if mat_id[p] < 0:
            lam = 0.28 * 190000.0 / ((1.28) * (1 - 0.56))
            mu = 190000.0 / (2 * 1.28)
            Dp = np.zeros((6, 6))
            Dp[:3, :3] = lam
            Dp[[0, 1, 2], [0, 1, 2]] += 2 * mu
            Dp[[3, 4, 5], [3, 4, 5]] = mu
            es[p], eD[p], et[:, p] = Dp @ eps[p], Dp, hist[:, p]
        else:
            s_, d_, st = port.j2_evaluate(mats[mat_id[p]], eps[p], st)
            es[p], eD[p], et[:, p] = s_, d_, st[13:]
            yielded += st[19] > st[6]
    assert yielded > 50
    assert rel_l2(sig, es) <= REL_K
    assert rel_l2(D, eD) <= REL_K
    for c in range(13):
        assert rel_l2(trial[c], et[c]) <= REL_STATE, c
