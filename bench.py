#!/usr/bin/env python3
"""Benchmark of the element-integration + global-assembly path on one MI355X.

Workload (default): the 64^3 C3D20 block — 262,144 elements, 1,085,825 nodes,
3,257,475 DOF, 557,286,921 CSR nonzeros — interior nodes jittered +-10% h (seed 0),
E_e ~ U[150000, 250000] (seed 1), nu = 0.3, smooth seeded u (seed 2). One step is
one full assembly of the CSR values of K (pattern reused, built once outside the
clock) plus the internal force f_int = K u, all FP64, inputs resident in HBM.

Prints one JSON line: throughput (elements/s), the dominant kernel's roofline figure
from HIP events recorded on the library's own stream, and the reference CPU assembly
timed on this host over a bounded sample.

  python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--n 64] [--family C3D20]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6    # TFLOP/s, MI355X dense FP64 (vector = matrix)
PEAK_HBM = 8000.0   # GB/s, MI355X HBM3E
J2_FLOP_PER_POINT = 300  # radial return, stress and compact tangent of a yielding point
ADJ_FWD_FLOP_C3D8 = 5586  # matrix-free r = K(x, rho) u per C3D8 element


def flops_per_element(nn: int, ngp: int, residual: bool) -> float:
    """Algorithmic FP64 count: Jacobian, gradients, sparse-B D*G_b and G_a^T (D G_b)
    over the upper node-pair blocks, scaling and the 3x3 inverse, per Gauss point;
    plus the matrix-free residual when requested."""
    k = ngp * (2 * (9 * nn + 9 * nn + 54 * nn + 27 * nn * (nn + 1) / 2) + 18 * nn + 50)
    r = ngp * 2 * (18 * nn + 36) if residual else 0
    return k + r


def bytes_per_assembly(nnz: int, n_nodes: int, n_elem: int, nn: int, residual: bool) -> float:
    """Algorithmic HBM bytes: K values, coordinates, connectivity, E per element
    (+ u read and f_int written for the residual)."""
    b = 8 * nnz + 24 * n_nodes + 4 * nn * n_elem + 8 * n_elem
    return b + (48 * n_nodes if residual else 0)


def load_traffic(workload: str, variant: int, key: str = "hbm_bytes_per_step"):
    """HBM bytes per assembly (or Newton iteration) measured by rocprofv3 --pmc
    (FETCH_SIZE, WRITE_SIZE), copy-calibrated, as committed in profiles/traffic.json;
    None if not measured for this variant."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("asm_variant") != variant:
        return None
    return rec.get(key)


def adjoint_ratio(reps: int = 20) -> dict:
    """Adjoint/forward cost on the 96x48x48 C3D8 SIMP block (jitter 5%, seeds 5-8):
    median over `reps` warm runs of the VJP kernel time over the forward r = K(x,rho)u
    kernel time, both from HIP events on the library's stream."""
    import statistics
    import numpy as np
    import torch
    from calculixpp_amd import meshgen
    from calculixpp_amd.autograd import ResidualMesh
    m, rho = meshgen.adjoint_block()
    mesh = ResidualMesh(m)
    dev = mesh.ctx.device
    x = torch.from_numpy(m.coords).to(dev)
    E = torch.from_numpy(m.E).to(dev)
    u = torch.from_numpy(meshgen.smooth_field(m.coords, 7)).to(dev)
    w = torch.from_numpy(np.random.default_rng(8).normal(size=(m.n_nodes, 3))).to(dev)
    r = torch.empty_like(u)
    dx, du, dE = torch.empty_like(x), torch.empty_like(u), torch.empty_like(E)
    mesh.forward_raw(x, E, u, r)
    mesh.vjp_raw(x, E, u, w, dx, dE, du)
    ctx = mesh.ctx
    fwd, vjp = [], []
    for _ in range(reps):
        for fn, acc in ((lambda: mesh.forward_raw(x, E, u, r), fwd),
                        (lambda: mesh.vjp_raw(x, E, u, w, dx, dE, du), vjp)):
            ctx.profile_reset()
            ctx.profile(True)
            fn()
            ctx.profile(False)
            acc.append(sum(v[1] for k, v in ctx.profile_stats().items()))
    f_ms, v_ms = statistics.median(fwd), statistics.median(vjp)
    # forward roofline: 5,586 FLOP per C3D8 element (Jacobian, gradients, strain, stress,
    # B^T sigma); bytes: coordinates, u and r per node, connectivity and E per element
    flop = ADJ_FWD_FLOP_C3D8 * m.n_elem
    byt = 72 * m.n_nodes + (4 * 8 + 8) * m.n_elem
    t_att = max(flop / (PEAK_FP64 * 1e12), byt / (PEAK_HBM * 1e9))
    return {"workload": "96x48x48 C3D8 SIMP block, r = K(x,rho)u and its VJP (dx, drho, du)",
            "elements": m.n_elem, "forward_ms": f_ms, "vjp_ms": v_ms, "ratio": v_ms / f_ms,
            "reps": reps,
            "forward_roofline": {"bound": "mfma" if flop / (PEAK_FP64 * 1e12) >= byt / (PEAK_HBM * 1e9)
                                 else "hbm", "flop": flop, "algorithmic_bytes": byt,
                                 "t_attainable_ms": round(t_att * 1e3, 4),
                                 "frac": round(t_att * 1e3 / f_ms, 4)}}


def j2_newton(shape=(128, 128, 64), reps: int = 5, target: float = 0.40,
              cpu_shape=(32, 32, 16), cpu: bool = True) -> dict:
    """One J2 Newton iteration (radial return at every Gauss point, consistent tangent
    K_t into the CSR values, internal force) on the 128x128x64 C3D8 block with seeded
    committed history (seed 3). The smooth displacement (seed 7) is scaled by bisection
    until the fraction of yielding points is within 0.005 of `target`."""
    import statistics
    import torch
    from calculixpp_amd import fem, meshgen
    dev = torch.device("cuda", 0)
    m = meshgen.j2_block(*shape)
    hist = meshgen.seeded_history(m.n_points, 3)
    a = fem.TangentAssembler(m)
    mp = fem.MaterialPoints(m.n_points, dev, committed=hist)
    base = torch.from_numpy(meshgen.smooth_field(m.coords, 7, 1.0)).to(dev)
    u = torch.empty_like(base)
    vals = a.new_values()
    f_int = torch.empty(3 * m.n_nodes, dtype=torch.float64, device=dev)

    def yield_frac(amp: float) -> float:
        torch.mul(base, amp, out=u)
        a.assemble_tangent(u, mp, vals, f_int)
        return float((mp.trial[6] > mp.committed[6]).double().mean())

    lo, hi = 1e-6, 1e-1
    amp, frac = hi, yield_frac(hi)
    for _ in range(60):
        amp = (lo * hi) ** 0.5
        frac = yield_frac(amp)
        if abs(frac - target) <= 0.005:
            break
        lo, hi = (amp, hi) if frac < target else (lo, amp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        a.assemble_tangent(u, mp, vals, f_int)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / reps
    ctx = a.ctx
    g = a.capture(u, mp, vals, f_int)  # the same iteration as one HIP graph
    g.launch()
    ctx.check_kernels()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.launch()
    ctx.check_kernels()
    graph_ms = (time.perf_counter() - t0) * 1e3 / reps
    graph_nodes = g.nodes
    del g
    per = []
    kern = {}
    for _ in range(reps):
        ctx.profile_reset()
        ctx.profile(True)
        a.assemble_tangent(u, mp, vals, f_int)
        ctx.profile(False)
        st = ctx.profile_stats()
        per.append(sum(v[1] for v in st.values()))
        for k, v in st.items():
            kern.setdefault(k, []).append(v[1])
    out = {"workload": f"{shape[0]}x{shape[1]}x{shape[2]} C3D8 J2 block, one Newton iteration "
                       "(return map + K_t + f_int)",
           "elements": m.n_elem, "gauss_points": m.n_points, "nnz": a.pattern.nnz,
           "amplitude": amp, "yield_fraction": frac, "ms_per_iteration": wall_ms,
           "graph_ms_per_iteration": graph_ms, "graph_nodes": graph_nodes,
           "kernel_ms": statistics.median(per),
           "kernels": {k: round(statistics.median(v), 4) for k, v in kern.items()},
           "elements_per_s": m.n_elem / (wall_ms * 1e-3)}
    # roofline of the iteration's launches: K_t + f_int per element and 300 FLOP per
    # Gauss point of return map; bytes as the assembly plus the history read and written
    flop = m.n_elem * flops_per_element(8, 8, True) + J2_FLOP_PER_POINT * m.n_points
    byt = bytes_per_assembly(a.pattern.nnz, m.n_nodes, m.n_elem, 8, True) + 208 * m.n_points
    t_att = max(flop / (PEAK_FP64 * 1e12), byt / (PEAK_HBM * 1e9))
    out["roofline"] = {"bound": "mfma" if flop / (PEAK_FP64 * 1e12) >= byt / (PEAK_HBM * 1e9)
                       else "hbm", "flop": flop, "algorithmic_bytes": byt,
                       "t_attainable_ms": round(t_att * 1e3, 4),
                       "frac": round(t_att * 1e3 / out["kernel_ms"], 4),
                       "traffic": load_traffic("J2-C3D8-128x128x64",
                                               ctx.options.get("asm_variant"),
                                               "hbm_bytes_per_iteration")}
    if cpu:
        from oracle.oracle import Oracle, aos_state, ref_available
        kind = "reference" if ref_available() else "port"
        cm = meshgen.j2_block(*cpu_shape)
        ch = meshgen.seeded_history(cm.n_points, 3)
        cu = meshgen.smooth_field(cm.coords, 7, amp)
        sec = Oracle(kind).time_material_tangent(cm, cu, aos_state(ch))
        out["cpu_baseline"] = {"value": cm.n_elem / sec, "unit": "elements/s", "cores": 1,
                               "kind": kind,
                               "sample": f"{cpu_shape[0]}x{cpu_shape[1]}x{cpu_shape[2]} C3D8 J2 "
                                         f"block, same seeds and amplitude, {sec:.2f} s"}
        out["gpu_over_cpu"] = out["elements_per_s"] / out["cpu_baseline"]["value"]
    return out


def batch_step(n_meshes: int = 512, reps: int = 3, cpu: bool = True, cpu_seconds: float = 15.0,
               log=None) -> dict:
    """The batched workload: n_meshes mixed C3D8/C3D20/C3D6 boxes with log-uniform
    element counts in [100, 50000] (seed 4), half of them J2, 4 load cases each. One
    step = the linear meshes' K plus f_int for the 4 cases, and for the J2 meshes per
    case the return map, K_t and f_int; elements of every mesh and family go through one
    persistent launch per assembly. Throughput counts element evaluations (a linear
    element once, a J2 element once per case)."""
    import torch
    from calculixpp_amd import autotune, meshgen
    from calculixpp_amd.batch import BatchStep
    from calculixpp_amd import fem
    t0 = time.perf_counter()
    recs = meshgen.batch_meshes(n_meshes)
    t1 = time.perf_counter()
    ctx = fem.context()
    step = BatchStep(recs, ctx, variant=5)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    if log:
        log(f"[bench] batch: {n_meshes} meshes generated in {t1 - t0:.1f}s, uploaded + "
            f"patterns in {t2 - t1:.1f}s")
    tune = autotune.tune_persistent(step.run, ctx, log=log)
    n_lin = step.lin.model.n_elem if step.lin else 0
    n_j2 = step.j2.model.n_elem if step.j2 else 0
    evals = n_lin + step.n_cases * n_j2
    step.run()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(reps):
        step.run()
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - ts) * 1e3 / reps
    step.capture()
    step.launch()
    ctx.check_kernels()
    ts = time.perf_counter()
    for _ in range(reps):
        step.launch()
    ctx.check_kernels()
    graph_ms = (time.perf_counter() - ts) * 1e3 / reps
    ctx.profile_reset()
    ctx.profile(True)
    step.launch()
    step.run()
    ctx.profile(False)
    kern = {k: {"launches": v[0], "ms": round(v[1], 4)} for k, v in ctx.profile_stats().items()}
    nnz = sum(a.pattern.nnz for a in [getattr(step, "lin_asm", None), getattr(step, "j2_asm", None)]
              if a is not None)
    out = {"workload": f"{n_meshes} mixed C3D8/C3D20/C3D6 meshes, log-uniform 100..50000 "
                       f"elements, half J2, {step.n_cases} load cases",
           "meshes": n_meshes, "elements": n_lin + n_j2, "linear_elements": n_lin,
           "j2_elements": n_j2, "element_evaluations": evals, "nnz": nnz,
           "ms_per_step": eager_ms, "graph_ms_per_step": graph_ms,
           "evaluations_per_s": evals / (min(eager_ms, graph_ms) * 1e-3),
           "autotune": tune, "kernels_one_graph_plus_one_eager_step": kern}
    ctx.set_option("asm_variant", 7)
    del step
    torch.cuda.empty_cache()
    if cpu:
        from oracle.oracle import Oracle, aos_state, ref_available
        kind = "reference" if ref_available() else "port"
        o = Oracle(kind)
        sec, ev, used = 0.0, 0, 0
        for r in recs:  # the batch's meshes in order, until about cpu_seconds of work
            if sec >= cpu_seconds:
                break
            used += 1
            m = r["model"]
            if r["j2"]:
                for u in r["u"]:
                    sec += o.time_material_tangent(m, u, aos_state(r["history"]))
                ev += len(r["u"]) * m.n_elem
            else:
                sec += o.time_linear_static(m, r["u"][0])
                ev += m.n_elem
        out["cpu_baseline"] = {"value": ev / sec, "unit": "element evaluations/s", "cores": 1,
                               "kind": kind, "sample": f"the batch's first {used} meshes "
                               f"({ev} element evaluations, {sec:.1f} s on one host thread)"}
        out["gpu_over_cpu"] = out["evaluations_per_s"] / out["cpu_baseline"]["value"]
    return out


def recover_leg(model, u, family: int, reps: int = 5, cpu_n: int = 16, cpu: bool = True) -> dict:
    """Field recovery (recover_fields: nodal stress and strain, f_int, reactions) of the
    displacement u on the assembly workload, from HIP events on the library's stream.
    Algorithmic counts per element: n_gp (72 n + 80) FLOP (Jacobian, gradients,
    displacement gradient and B^T sigma at 9n FMA each, inverse, strain and stress);
    bytes 192 N_nodes (coords, u, f_ext in; strain, stress, reaction out) +
    (4n + 16) N_e (connectivity, E, nu)."""
    import statistics
    import torch
    from calculixpp_amd import recover
    from calculixpp_amd.model import GAUSS, NODES
    r = recover.FieldRecovery(model)
    ctx = r.ctx
    r.run(u)
    times = []
    for _ in range(reps):
        ctx.profile_reset()
        ctx.profile(True)
        r.run(u)
        ctx.profile(False)
        times.append(sum(v[1] for v in ctx.profile_stats().values()))
    stats = ctx.profile_stats()
    ms = statistics.median(times)
    nn, ngp = NODES[family], GAUSS[family]
    fpe = ngp * (72 * nn + 80)
    flop = fpe * model.n_elem
    byt = 192 * model.n_nodes + (4 * nn + 16) * model.n_elem
    t_att = max(flop / (PEAK_FP64 * 1e12), byt / (PEAK_HBM * 1e9))
    out = {"workload": "recover_fields of a smooth displacement on the assembly block",
           "elements": model.n_elem, "nodes": model.n_nodes, "ms": ms,
           "elements_per_s": model.n_elem / (ms * 1e-3),
           "kernels_ms": {k: round(v[1], 4) for k, v in stats.items()},
           "flop_per_element": fpe, "algorithmic_bytes": byt,
           "roofline": {"bound": "mfma" if flop / (PEAK_FP64 * 1e12) >= byt / (PEAK_HBM * 1e9)
                        else "hbm", "t_attainable_ms": round(t_att * 1e3, 4),
                        "frac": round(t_att / (ms * 1e-3), 4)}}
    del r
    torch.cuda.synchronize()
    if cpu:
        from calculixpp_amd import meshgen
        from oracle.oracle import Oracle, ref_available
        kind = "reference" if ref_available() else "port"
        cm = meshgen.block(cpu_n, family)
        cu = meshgen.smooth_field(cm.coords, 2)
        o = Oracle(kind)
        t0 = time.perf_counter()
        o.recover_fields(cm, cu)
        sec = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cm.n_elem / sec, "unit": "elements/s", "cores": 1,
                               "kind": kind, "sample": f"{cpu_n}^3 block of the same family "
                               f"and seeds ({cm.n_elem} elements, recover_fields, {sec:.1f} s "
                               "on one host thread)"}
        out["gpu_over_cpu"] = out["elements_per_s"] / out["cpu_baseline"]["value"]
    return out


def pcg_leg(n: int = 64, iters: int = 96) -> dict:
    """The device solve (cxg_pcg, Jacobi-preconditioned CG) on the n^3 C3D20 block of the
    assembly workload, clamped on x = 0 with a z load on x = 1: `iters` iterations (a
    multiple of the 16 between host residual checks) timed by HIP events. Bytes per
    iteration: the CSR product (8 + 4 per nonzero, 8 per row) and the vector work (x, r,
    z, p, q, 1/diag: 128 per row). No CPU baseline: the reference solves with SciPP,
    which is not available."""
    import numpy as np
    import torch
    from calculixpp_amd import _lib, fem, meshgen, solve
    from calculixpp_amd.model import Model
    b = meshgen.block(n, _lib.C3D20)
    x0 = np.nonzero(b.coords[:, 0] < 1e-12)[0].astype(np.int32)
    x1 = np.nonzero(b.coords[:, 0] > 1.0 - 1e-12)[0].astype(np.int32)
    m = Model(b.coords, b.blocks, b.E, b.nu, spc_node=np.repeat(x0, 3),
              spc_comp=np.tile(np.arange(1, 4, dtype=np.int32), len(x0)),
              spc_value=np.zeros(3 * len(x0)), cload_node=x1,
              cload_comp=np.full(len(x1), 3, np.int32), cload_value=np.full(len(x1), -1e-3))
    ctx = fem.context()
    a = fem.Assembler(m, ctx)
    vals = a.new_values()
    a.assemble(vals)
    rhs = a.rhs()
    nnz, nf = a.pattern.nnz, a.n_free
    rp, ci, va = a.pattern.row_ptr, a.pattern.col_idx, vals[:nnz]
    torch.cuda.synchronize()
    ctx.profile_reset()
    ctx.profile(True)
    rel = None
    try:
        _, _, rel = solve.pcg(ctx, rp, ci, va, rhs, tol=0.0, max_iter=iters)
    except _lib.CxgError as e:  # tol 0 is never reached: x holds the last iterate
        if "did not reach" not in str(e):
            raise
    ctx.profile(False)
    st = ctx.profile_stats()
    ms_iter = st["pcg_iterations"][1] / iters
    byt = 12 * nnz + 8 * (nf + 1) + 128 * nf
    t_att = byt / (PEAK_HBM * 1e9)
    out = {"workload": f"{n}^3 C3D20 block clamped on x = 0, z load on x = 1: Jacobi-PCG",
           "n_free": nf, "nnz": nnz, "iterations": iters, "ms_per_iteration": round(ms_iter, 4),
           "setup_ms": round(st["pcg_setup"][1], 4),
           "roofline": {"bound": "hbm", "algorithmic_bytes_per_iteration": byt,
                        "t_attainable_ms": round(t_att * 1e3, 4),
                        "frac": round(t_att * 1e3 / ms_iter, 4),
                        "achieved_GBps": round(byt / (ms_iter * 1e-3) / 1e9, 1)},
           "cpu_baseline": None,
           "note": "the reference's solve (SciPP spsolve / cg) is absent: no CPU baseline"}
    del a, vals
    torch.cuda.empty_cache()
    return out


def mass_geometric_leg(asm, family: int, reps: int = 3, cpu_n: int = 16, cpu: bool = True) -> dict:
    """Consistent mass, lumped mass and geometric stiffness into the CSR values of the
    assembly workload's pattern (density 7.85e-9, seeded Gauss-point stress), from HIP
    events on the library's stream. Both matrices are k_ab I3 blocks written through
    the same element arena and row gather as K; algorithmic bytes are the CSR values
    (8 nnz) plus coordinates, connectivity and one value per element or point."""
    import statistics
    import numpy as np
    import torch
    from calculixpp_amd.model import GAUSS, NODES
    dev = asm.ctx.device
    model = asm.model
    rho = torch.full((model.n_elem,), 7.85e-9, dtype=torch.float64, device=dev)
    gs = torch.from_numpy(np.random.default_rng(45).normal(scale=100.0, size=(model.n_points, 6))).to(dev)
    vals = asm.new_values()
    ctx = asm.ctx
    runs = {"mass": lambda: asm.assemble_mass(vals, False, rho),
            "mass_lumped": lambda: asm.assemble_mass(vals, True, rho),
            "geometric": lambda: asm.assemble_geometric(vals, gs)}
    nn, ngp = NODES[family], GAUSS[family]
    nnz = asm.pattern.nnz
    out = {"workload": "mass (consistent, lumped) and geometric stiffness on the assembly "
                       "block's CSR pattern", "elements": model.n_elem, "nnz": nnz}
    for key, fn in runs.items():
        fn()
        times = []
        for _ in range(reps):
            ctx.profile_reset()
            ctx.profile(True)
            fn()
            ctx.profile(False)
            times.append(sum(v[1] for v in ctx.profile_stats().values()))
        ms = statistics.median(times)
        extra = 48 * ngp * model.n_elem if key == "geometric" else 8 * model.n_elem
        byt = 8 * nnz + 24 * model.n_nodes + 4 * nn * model.n_elem + extra
        t_att = byt / (PEAK_HBM * 1e9)
        out[key] = {"ms": ms, "elements_per_s": model.n_elem / (ms * 1e-3),
                    "kernels_ms": {k: round(v[1], 4) for k, v in ctx.profile_stats().items()},
                    "roofline": {"bound": "hbm", "algorithmic_bytes": byt,
                                 "t_attainable_ms": round(t_att * 1e3, 4),
                                 "frac": round(t_att / (ms * 1e-3), 4)}}
    del vals
    torch.cuda.synchronize()
    if cpu:
        from calculixpp_amd import meshgen
        from calculixpp_amd.model import Model
        from oracle.oracle import Oracle, ref_available
        kind = "reference" if ref_available() else "port"
        b = meshgen.block(cpu_n, family)
        cm = Model(b.coords, b.blocks, b.E, b.nu, rho=7.85e-9)
        o = Oracle(kind)
        t0 = time.perf_counter()
        o.assemble_mass(cm)
        sec = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cm.n_elem / sec, "unit": "elements/s", "cores": 1,
                               "kind": kind, "sample": f"{cpu_n}^3 block of the same family "
                               f"({cm.n_elem} elements, assemble_mass, {sec:.1f} s on one "
                               "host thread)"}
        out["gpu_over_cpu"] = out["mass"]["elements_per_s"] / out["cpu_baseline"]["value"]
    return out


def cpu_baseline(family: int, n: int, with_residual: bool) -> dict:
    """Reference CPU assembly (+ internal force) on an n^3 sample of the same
    workload, one thread (the reference is single-threaded)."""
    from calculixpp_amd import meshgen
    from oracle.oracle import Oracle, ref_available
    kind = "reference" if ref_available() else "port"
    model = meshgen.block(n, family)
    u = meshgen.smooth_field(model.coords, 2) if with_residual else None
    sec = Oracle(kind).time_linear_static(model, u)
    return {"value": model.n_elem / sec, "unit": "elements/s", "cores": 1, "kind": kind,
            "sample": f"{n}^3 block of the same family and seeds ({model.n_elem} elements, "
                      f"assemble_linear_static{' + internal_force' if with_residual else ''}, "
                      f"{sec:.1f} s on one host thread)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--family", default="C3D20")
    ap.add_argument("--cpu-n", type=int, default=16, help="edge of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-residual", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adj-reps", type=int, default=20)
    ap.add_argument("--no-j2", action="store_true", help="skip the J2 Newton-iteration leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-meshes leg")
    ap.add_argument("--no-recover", action="store_true", help="skip the field-recovery leg")
    ap.add_argument("--no-mass", action="store_true",
                    help="skip the mass / geometric-stiffness leg")
    ap.add_argument("--no-pcg", action="store_true", help="skip the device-solve leg")
    ap.add_argument("--batch-meshes", type=int, default=512)
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="extra cxg_set_option for the library context (repeatable)")
    ap.add_argument("--only", choices=["j2", "batch", "adjoint", "pcg"], default=None,
                    help="run one secondary leg alone and print its JSON (profiling)")
    ap.add_argument("--variant", type=int, default=0, help="force asm_variant (0 = autotune)")
    ap.add_argument("--patch", type=int, default=0, help="force patch_size (0 = autotune)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n * 8, "launches": 3}), flush=True)
        return 0
    if args.only:
        import torch
        torch.cuda.set_device(0)
        if args.option or args.variant:
            from calculixpp_amd import fem
            if args.variant:
                fem.context().set_option("asm_variant", args.variant)
            for opt in args.option:
                key, val = opt.split("=", 1)
                fem.context().set_option(key, int(val))
        if args.only == "j2":
            res = j2_newton(reps=args.steps, cpu=not args.no_cpu)
        elif args.only == "pcg":
            res = pcg_leg(args.n)
        elif args.only == "batch":
            res = batch_step(args.batch_meshes, reps=args.steps, cpu=not args.no_cpu,
                             log=lambda m: print(m, file=sys.stderr, flush=True))
        else:
            res = adjoint_ratio(args.adj_reps)
        print(json.dumps(res), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding of this path is a design, not a build", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from calculixpp_amd import _lib, fem, meshgen
    from calculixpp_amd.model import GAUSS, NODES

    family = _lib.TYPE_NAMES[args.family]
    residual = not args.no_residual
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    t0 = time.perf_counter()
    model = meshgen.block(args.n, family)
    u_host = meshgen.smooth_field(model.coords, 2)
    t1 = time.perf_counter()
    u = torch.from_numpy(u_host).to(dev) if residual else None
    from calculixpp_amd import autotune
    if args.variant and args.patch:
        choice = {"asm_variant": args.variant, "patch_size": args.patch, "trials": []}
    else:
        choice = autotune.tune_linear(model, u, cache_path=None, log=lambda m: print(
            m, file=sys.stderr, flush=True))
    t2 = time.perf_counter()
    for opt in args.option:
        key, val = opt.split("=", 1)
        fem.context().set_option(key, int(val))
    asm = autotune.make_assembler(model, choice["asm_variant"], choice["patch_size"])
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    print(f"[bench] mesh {t1 - t0:.1f}s, tune {t2 - t1:.1f}s, upload+colour+pattern "
          f"{t3 - t2:.1f}s: {model.n_elem} elements, {model.n_nodes} nodes, nnz "
          f"{asm.pattern.nnz}, {asm.mesh.n_colors} colours, variant {choice['asm_variant']}, "
          f"patch {choice['patch_size']}", file=sys.stderr, flush=True)
    vals = asm.new_values()
    f_int = torch.empty(3 * model.n_nodes, dtype=torch.float64, device=dev) if residual else None
    asm.assemble(vals, u, f_int)  # set-up: arenas allocated, gather order built

    for _ in range(args.warmup):
        asm.assemble(vals, u, f_int)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(args.steps):
        asm.assemble(vals, u, f_int)
    torch.cuda.synchronize()
    te = time.perf_counter()
    ms_step = (te - ts) * 1e3 / args.steps
    value = model.n_elem / (ms_step * 1e-3)
    print(f"[bench] {ms_step:.3f} ms/step, {value:.4g} elements/s", file=sys.stderr, flush=True)

    # Per-kernel durations from HIP events on the library's stream. One assembly is
    # the element kernel followed by the row-gather kernel (two launches, one step);
    # the roofline fraction is max(F/peak_fp64, B/peak_hbm) / t over both launches.
    ctx = asm.ctx
    ctx.profile_reset()
    ctx.profile(True)
    reps = 3
    for _ in range(reps):
        asm.assemble(vals, u, f_int)
    ctx.profile(False)
    stats = ctx.profile_stats()
    kernels = {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 4),
                   "ms_per_step": round(v[1] / reps, 4), "elements": v[2]}
               for k, v in stats.items()}
    step_ms = sum(v[1] for v in stats.values()) / reps
    name = max(stats, key=lambda k: stats[k][1])
    nn, ngp = NODES[family], GAUSS[family]
    fpe = flops_per_element(nn, ngp, residual)
    flop = fpe * model.n_elem
    bpa = bytes_per_assembly(asm.pattern.nnz, model.n_nodes, model.n_elem, nn, residual)
    t_flop, t_byte = flop / (PEAK_FP64 * 1e12), bpa / (PEAK_HBM * 1e9)
    bound = "mfma" if t_flop >= t_byte else "hbm"
    achieved_tf = flop / (step_ms * 1e-3) / 1e12
    achieved_gb = bpa / (step_ms * 1e-3) / 1e9
    traffic = load_traffic(f"{args.family}-{args.n}", choice["asm_variant"])
    roofline = {"bound": bound,
                "achieved": round(achieved_tf if bound == "mfma" else achieved_gb, 3),
                "peak": PEAK_FP64 if bound == "mfma" else PEAK_HBM,
                "unit": "TFLOP/s" if bound == "mfma" else "GB/s",
                "frac": round(max(t_flop, t_byte) / (step_ms * 1e-3), 4),
                "traffic": traffic,
                "kernel": " + ".join(sorted(stats)), "dominant": name,
                "step_kernel_ms": round(step_ms, 4),
                "flop_per_element": fpe, "flop_per_step": flop,
                "algorithmic_bytes_per_step": bpa,
                "hbm_algorithmic_GBps": round(achieved_gb, 1),
                "t_attainable_ms": round(max(t_flop, t_byte) * 1e3, 4)}

    adjoint = None if args.no_adjoint else adjoint_ratio(args.adj_reps)
    if adjoint:
        print(f"[bench] adjoint/forward {adjoint['ratio']:.3f} ({adjoint['vjp_ms']:.4f} / "
              f"{adjoint['forward_ms']:.4f} ms)", file=sys.stderr, flush=True)

    j2 = None if args.no_j2 else j2_newton(cpu=not args.no_cpu)
    if j2:
        print(f"[bench] J2 Newton iteration {j2['ms_per_iteration']:.3f} ms "
              f"({j2['elements_per_s']:.4g} elements/s, yield {j2['yield_fraction']:.3f})",
              file=sys.stderr, flush=True)

    rec = None
    if not args.no_recover and u is not None:
        rec = recover_leg(model, u.reshape(-1, 3), family, cpu_n=args.cpu_n, cpu=not args.no_cpu)
        print(f"[bench] recover_fields {rec['ms']:.3f} ms ({rec['elements_per_s']:.4g} "
              f"elements/s)", file=sys.stderr, flush=True)

    mass = None if args.no_mass else mass_geometric_leg(asm, family, cpu_n=args.cpu_n,
                                                        cpu=not args.no_cpu)
    if mass:
        print(f"[bench] mass {mass['mass']['ms']:.3f} ms, lumped {mass['mass_lumped']['ms']:.3f} "
              f"ms, geometric {mass['geometric']['ms']:.3f} ms", file=sys.stderr, flush=True)

    pcg = None if args.no_pcg else pcg_leg(args.n)
    if pcg:
        print(f"[bench] pcg {pcg['ms_per_iteration']:.3f} ms per iteration "
              f"({pcg['roofline']['achieved_GBps']:.0f} GB/s)", file=sys.stderr, flush=True)

    batch = None if args.no_batch else batch_step(
        args.batch_meshes, cpu=not args.no_cpu, log=lambda m: print(m, file=sys.stderr, flush=True))
    if batch:
        print(f"[bench] batch step {batch['ms_per_step']:.2f} ms eager, "
              f"{batch['graph_ms_per_step']:.2f} ms graph ({batch['evaluations_per_s']:.4g} "
              "element evaluations/s)", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(family, args.cpu_n, residual)
        print(f"[bench] cpu baseline {cpu['value']:.4g} elements/s ({cpu['kind']})",
              file=sys.stderr, flush=True)

    out = {
        "metric": "C3D20 K-assembly elements/s + % FP64/HBM roofline; adjoint/forward cost",
        "value": value, "unit": "elements/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.n}^3 {args.family} jittered block: CSR K values"
                               f"{' + internal force' if residual else ''}",
                   "elements": model.n_elem, "nodes": model.n_nodes,
                   "dof": 3 * model.n_nodes, "nnz": asm.pattern.nnz,
                   "colours": asm.mesh.n_colors, "asm_variant": choice["asm_variant"],
                   "patch_size": choice["patch_size"]},
        "autotune": choice["trials"],
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "adjoint": adjoint,
        "j2_newton": j2,
        "recover": rec,
        "mass_geometric": mass,
        "pcg": pcg,
        "batch": batch,
        "gpu_over_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
