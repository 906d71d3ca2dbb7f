"""The allocation-free internal force (cxg_internal_force_noalloc) against the allocating
entry point it mirrors (cxg_internal_force), both called through the C ABI.

The two share their kernels; the allocation-free one puts the element forces into the
library's fixed device scratch array instead of an arena of the mesh. Outputs must be
bit-identical on two calls in a row (the second on a smaller mesh, so the scratch array
still holds the first call's larger contents), and within rel-L2 1e-12 of the oracle's
fem::internal_force restatement (src/fem/stress.cpp:125-145). A mesh larger than the
scratch array is refused with CXG_ERR_ARG before anything is enqueued.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 8 * 1024 * 1024


def rel_l2(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def fem():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from calculixpp_amd import fem
    return fem


def _both(fem, model, seed):
    """f_int of cxg_internal_force and of cxg_internal_force_noalloc on the same inputs."""
    import torch
    from calculixpp_amd import meshgen
    mesh = fem.DeviceMesh(model)
    ctx = mesh.ctx
    u = torch.from_numpy(meshgen.smooth_field(model.coords, seed)).to(ctx.device)
    f_ref = torch.full((3 * model.n_nodes,), np.nan, dtype=torch.float64, device=ctx.device)
    f_new = torch.full_like(f_ref, np.nan)
    p = fem._ptr
    ctx.check(ctx.lib.cxg_internal_force(ctx.h, mesh.h, p(mesh.coords), p(mesh.E), p(mesh.nu),
                                         p(u), p(f_ref)))
    ctx.check(ctx.lib.cxg_internal_force_noalloc(ctx.h, mesh.h, p(mesh.coords), p(mesh.E),
                                                 p(mesh.nu), p(u), p(f_new)))
    torch.cuda.synchronize()
    return u.cpu().numpy(), f_ref.cpu().numpy(), f_new.cpu().numpy()


def test_noalloc_matches_allocating_entry_point_twice(fem, port):
    from calculixpp_amd import _lib, meshgen
    from calculixpp_amd.model import Model
    # first call: a mixed C3D20 + C3D8 mesh; second call: a smaller C3D6 + C3D8 mesh
    big = meshgen.block(12, _lib.C3D20, seed_geo=41, seed_E=42)
    coords, hexes = meshgen.hex_grid(5, 4, 3, (1.0, 0.8, 0.6), _lib.C3D8)
    coords = meshgen.jitter(coords, 0.2, 0.1, 43, (1.0, 0.8, 0.6))
    wed = np.concatenate([hexes[30:][:, [0, 1, 2, 4, 5, 6]], hexes[30:][:, [0, 2, 3, 4, 6, 7]]])
    small = Model(coords, [(_lib.C3D8, hexes[:30]), (_lib.C3D6, wed)],
                  np.linspace(1.5e5, 2.5e5, 30 + len(wed)), 0.3)
    for model, seed in ((big, 44), (small, 45)):
        u, f_ref, f_new = _both(fem, model, seed)
        assert np.isfinite(f_new).all()
        assert np.array_equal(f_new, f_ref)
        assert rel_l2(f_new, port.internal_force(model, u)) <= 1e-12


def test_noalloc_refuses_a_mesh_beyond_the_scratch(fem):
    import torch
    from calculixpp_amd import _lib, meshgen
    from calculixpp_amd.model import Model
    coords, conn = meshgen.hex_grid(71, 71, 70, (1.0, 1.0, 70 / 71), _lib.C3D8)
    assert 3 * conn.size > SCRATCH_DOUBLES
    model = Model(coords, [(_lib.C3D8, conn)], 210000.0, 0.3)
    mesh = fem.DeviceMesh(model)
    ctx = mesh.ctx
    u = torch.zeros(3 * model.n_nodes, dtype=torch.float64, device=ctx.device)
    f = torch.empty_like(u)
    p = fem._ptr
    st = ctx.lib.cxg_internal_force_noalloc(ctx.h, mesh.h, p(mesh.coords), p(mesh.E), p(mesh.nu),
                                            p(u), p(f))
    assert st == _lib.ERR_ARG
    assert b"maximum is 8388608" in ctx.lib.cxg_last_error(ctx.h)
    # a null argument is refused the same way
    assert ctx.lib.cxg_internal_force_noalloc(ctx.h, mesh.h, None, p(mesh.E), p(mesh.nu),
                                              p(u), p(f)) == _lib.ERR_ARG
