"""Batches of uneven meshes with mixed families: one persistent launch (asm_variant 5) or
one launch per family (asm_variant 4 full rows, 7 upper blocks in the per-mesh gather order).

A small instance of the batched workload (meshgen.batch_meshes: mixed C3D8 / C3D20 /
C3D6 boxes, per-mesh E and nu, half of the meshes J2 with seeded history, 4 load
cases) runs through calculixpp_amd.batch.BatchStep; every mesh's CSR pattern, K or
K_t, internal force per load case and trial history are compared with the oracle run
on that mesh alone:
  CSR row_ptr / col_idx   bit-exact
  K, f_int                rel-L2 <= 1e-12
  trial history           rel-L2 <= 1e-10
The work-item size and the persistent grid size are swept; graph replays of the whole
step are compared bitwise with the eager step, and the variants bitwise with each other
(every entry is summed in the same element order whatever the launch shape).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL_K = 1e-12
REL_STATE = 1e-10


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from calculixpp_amd import fem
    return fem


@pytest.fixture(scope="module")
def records():
    from calculixpp_amd import meshgen
    return meshgen.batch_meshes(n_meshes=8, seed=11, e_min=20, e_max=300)


@pytest.fixture(scope="module")
def expected(records, port):
    from oracle.oracle import aos_state
    out = []
    for r in records:
        m = r["model"]
        if r["j2"]:
            cases = [port.assemble_material_tangent(m, u, aos_state(r["history"])) for u in r["u"]]
            out.append(("j2", cases))
        else:
            sys_ = port.assemble_linear_static(m)
            out.append(("lin", sys_, [port.internal_force(m, u) for u in r["u"]]))
    return out


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _check(step, records, expected):
    for k, i in enumerate(step.lin_idx):
        _, sys_e, f_e = expected[i]
        rp, ci, v = sys_e.csr()
        grp, gci, gv = step.lin.csr(k, step.lin_asm.pattern.row_ptr, step.lin_asm.pattern.col_idx,
                                    step.lin_vals)
        assert np.array_equal(grp, rp) and np.array_equal(gci, ci), i
        assert rel_l2(gv, v) <= REL_K, i
        for c in range(step.n_cases):
            assert rel_l2(step.lin.nodal(k, step.lin_f[c]), f_e[c]) <= REL_K, (i, c)
    for k, i in enumerate(step.j2_idx):
        _, cases = expected[i]
        for c, (sys_e, f_e, st_e) in enumerate(cases):
            rp, ci, v = sys_e.csr()
            grp, gci, gv = step.j2.csr(k, step.j2_asm.pattern.row_ptr, step.j2_asm.pattern.col_idx,
                                       step.j2_vals[c])
            assert np.array_equal(grp, rp) and np.array_equal(gci, ci), (i, c)
            assert rel_l2(gv, v) <= REL_K, (i, c)
            assert rel_l2(step.j2.nodal(k, step.j2_f[c]), f_e) <= REL_K, (i, c)
            trial = step.j2.points(k, step.j2_mp[c].trial)
            for q in range(13):
                assert rel_l2(trial[q], st_e[:, 13 + q]) <= REL_STATE, (i, c, q)


def test_batch_mix(records):
    fams = {t for r in records for t, _ in r["model"].blocks}
    assert len(fams) == 3
    assert 0 < sum(r["j2"] for r in records) < len(records)


@pytest.mark.parametrize("chunk,waves", [(1, 8), (3, 1), (16, 8), (64, 2)])
def test_persistent_batch_matches_oracle(gpu, records, expected, chunk, waves):
    from calculixpp_amd.batch import BatchStep
    ctx = gpu.context()
    ctx.set_option("chunk_elems", chunk)
    ctx.set_option("persistent_waves", waves)
    try:
        step = BatchStep(records, ctx, variant=5)
        step.run()
        _check(step, records, expected)
    finally:
        ctx.set_option("asm_variant", 7)
        ctx.set_option("chunk_elems", 4)
        ctx.set_option("persistent_waves", 8)


@pytest.mark.parametrize("variant", [4, 7])
def test_batch_per_family_kernels_match_oracle(gpu, records, expected, variant):
    from calculixpp_amd.batch import BatchStep
    ctx = gpu.context()
    try:
        step = BatchStep(records, ctx, variant=variant)
        step.run()
        _check(step, records, expected)
    finally:
        ctx.set_option("asm_variant", 7)


def _outputs(step):
    import torch
    torch.cuda.synchronize()
    return ([step.lin_vals.clone(), step.lin_f.clone()] +
            [x.clone() for v, f, mp in zip(step.j2_vals, step.j2_f, step.j2_mp) for x in (v, f, mp.trial)])


def test_batch_variants_same_bits(gpu, records):
    """Persistent (5), per-family full rows (4) and per-family upper blocks in the
    per-mesh gather order (7, also in node order with gather_tile 0) give the same bits."""
    import torch
    from calculixpp_amd.batch import BatchStep
    ctx = gpu.context()
    try:
        outs = []
        for variant, tile in [(5, 8), (4, 8), (7, 8), (7, 0)]:
            ctx.set_option("gather_tile", tile)
            step = BatchStep(records, ctx, variant=variant)
            step.run()
            outs.append(_outputs(step))
            del step
        for o in outs[1:]:
            assert all(torch.equal(a, b) for a, b in zip(outs[0], o))
    finally:
        ctx.set_option("asm_variant", 7)
        ctx.set_option("gather_tile", 8)


def test_tune_batch_choice_matches_oracle(gpu, records, expected):
    from calculixpp_amd import autotune
    from calculixpp_amd.batch import BatchStep
    ctx = gpu.context()
    try:
        step = BatchStep(records, ctx)
        best = autotune.tune_batch(step.run, ctx, candidates=[(4, 8), (16, 2)], steps=1)
        assert best["asm_variant"] in (5, 7) and len(best["trials"]) == 3
        assert ctx.options["asm_variant"] == best["asm_variant"]
        step.run()
        _check(step, records, expected)
    finally:
        ctx.set_option("asm_variant", 7)
        ctx.set_option("chunk_elems", 4)
        ctx.set_option("persistent_waves", 8)


@pytest.mark.parametrize("variant", [5, 7])
def test_batch_graph_replays_match_eager(gpu, records, variant):
    import torch
    from calculixpp_amd.batch import BatchStep
    ctx = gpu.context()
    try:
        step = BatchStep(records, ctx, variant=variant)
        step.run()
        torch.cuda.synchronize()
        ref_lin = (step.lin_vals.clone(), step.lin_f.clone())
        ref_j2 = [(v.clone(), f.clone(), mp.trial.clone())
                  for v, f, mp in zip(step.j2_vals, step.j2_f, step.j2_mp)]
        step.capture()
        for rep in range(3):
            step.lin_vals.fill_(-1.0)
            for v, mp in zip(step.j2_vals, step.j2_mp):
                v.fill_(-1.0)
                mp.trial.fill_(-1.0)
            step.launch()
            ctx.check_kernels()
            assert torch.equal(step.lin_vals, ref_lin[0]) and torch.equal(step.lin_f, ref_lin[1])
            for (v, f, t), v2, f2, mp in zip(ref_j2, step.j2_vals, step.j2_f, step.j2_mp):
                assert torch.equal(v, v2) and torch.equal(f, f2) and torch.equal(t, mp.trial)
            _ = (step.lin_f * 2).sum().item()
    finally:
        ctx.set_option("asm_variant", 7)
