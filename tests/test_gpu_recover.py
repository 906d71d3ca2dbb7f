"""Field recovery on the device (cxg_recover_fields via calculixpp_amd.recover) against
the oracle's recover_fields, which is pinned bit for bit to the reference core
(tests/test_oracle.py). Same seeded inputs; strain, stress, reactions and the
Gauss-point stress within rel-L2 1e-12 (the device sums B^T sigma where the reference
forms Ke ue, and its Gauss-point sums run in a different order).
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
def rec():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from calculixpp_amd import recover
    return recover


def _families():
    from calculixpp_amd import _lib
    return [_lib.C3D8, _lib.C3D20, _lib.C3D20R, _lib.C3D6]


def _check(got: dict, exp: dict, keys):
    for k in keys:
        assert rel_l2(got[k], exp[k]) <= REL, (k, rel_l2(got[k], exp[k]))


@pytest.mark.parametrize("fam_i", range(4))
def test_recover_fields_elastic(rec, port, fam_i):
    import torch
    from calculixpp_amd import meshgen
    m = meshgen.block(4, _families()[fam_i], seed_geo=51, seed_E=52, ny=3, nz=2)
    u = meshgen.smooth_field(m.coords, 53)
    r = rec.FieldRecovery(m)
    o = {k: v.cpu().numpy() for k, v in r.run(torch.from_numpy(u), gauss_stress=True).items()}
    exp = port.recover_fields(m, u, gauss_stress=True)
    _check(o, exp, ("strain", "stress", "reaction", "f_int", "gauss_stress"))


@pytest.mark.parametrize("fam_i", range(4))
def test_recover_fields_material_points(rec, port, fam_i):
    """The material-point form: J2 elements (radial return from the committed
    history, about half of the points yielding) mixed with elastic ones."""
    from calculixpp_amd import fem, meshgen
    from oracle.oracle import aos_state
    m = meshgen.j2_block(4, 3, 3, (1.0, 0.8, 0.6), family=_families()[fam_i], elastic_every=5,
                         jitter_frac=0.1)
    hist = meshgen.seeded_history(m.n_points, 61)
    u = meshgen.smooth_field(m.coords, 62, 1.5e-4)
    mp = fem.MaterialPoints(m.n_points, committed=hist)
    f = rec.recover_fields(m, u, mp)
    exp = port.recover_fields(m, u, aos_state(hist))
    _check({"stress": f.stress, "strain": f.strain, "reaction": f.reaction}, exp,
           ("stress", "strain", "reaction"))
    # the committed history is read, never advanced
    assert np.array_equal(mp.committed.cpu().numpy(), hist)


def test_recover_fields_loads_and_removed_elements(rec, port):
    """The loaded, clamped cantilever (RF = f_int - f_ext), then with every third
    element removed (*MODEL CHANGE): removed elements add nothing and their Gauss
    points are empty in recover_gauss_stress."""
    from calculixpp_amd import meshgen
    m = meshgen.cantilever()
    u = meshgen.smooth_field(m.coords, 71)
    f = rec.recover_fields(m, u)
    exp = port.recover_fields(m, u)
    _check({"stress": f.stress, "strain": f.strain, "reaction": f.reaction}, exp,
           ("stress", "strain", "reaction"))
    m.active = np.ones(m.n_elem, bool)
    m.active[::3] = False
    f = rec.recover_fields(m, u)
    exp = port.recover_fields(m, u, gauss_stress=True)
    _check({"stress": f.stress, "strain": f.strain, "reaction": f.reaction}, exp,
           ("stress", "strain", "reaction"))
    gs = rec.recover_gauss_stress(m, u)
    assert len(gs) == m.n_elem
    assert all(gs[e].shape == (0, 6) for e in range(0, m.n_elem, 3))
    got = np.concatenate([g if g.shape[0] else np.zeros((8, 6)) for g in gs])
    assert rel_l2(got, exp["gauss_stress"]) <= REL


def test_recover_mixed_families_one_mesh(rec, port):
    import torch
    from calculixpp_amd import meshgen
    from calculixpp_amd.model import Model
    c, blocks = meshgen.mixed_box(120, seed=81)
    n_elem = sum(b.shape[0] for _, b in blocks)
    m = Model(c, blocks, np.random.default_rng(83).uniform(1e5, 2e5, n_elem), 0.3)
    u = meshgen.smooth_field(m.coords, 82)
    o = {k: v.cpu().numpy() for k, v in
         rec.FieldRecovery(m).run(torch.from_numpy(u), gauss_stress=True).items()}
    exp = port.recover_fields(m, u, gauss_stress=True)
    _check(o, exp, ("strain", "stress", "reaction", "f_int", "gauss_stress"))


def test_recover_deterministic(rec):
    import torch
    from calculixpp_amd import _lib, meshgen
    m = meshgen.block(4, _lib.C3D20, seed_geo=91, seed_E=92)
    u = torch.from_numpy(meshgen.smooth_field(m.coords, 93))
    r = rec.FieldRecovery(m)
    a, b = r.run(u, gauss_stress=True), r.run(u, gauss_stress=True)
    for k in a:
        assert torch.equal(a[k], b[k]), k
