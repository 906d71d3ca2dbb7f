#!/usr/bin/env python3
"""Benchmark of the element-integration + global-assembly path on one MI355X.

Workload (default): the 64^3 C3D20 block — 262,144 elements, 1,085,825 nodes,
3,257,475 DOF, 557,286,921 CSR nonzeros — interior nodes jittered +-10% h (seed 0),
E_e ~ U[150000, 250000] (seed 1), nu = 0.3, smooth seeded u (seed 2). One step is
one full assembly of the CSR values of K (pattern reused, built once outside the
clock) plus the internal force f_int = K u, all FP64, inputs resident in HBM. The
same block of C3D20R elements (8-point rule) is a second assembly leg.

Prints one JSON line: throughput (elements/s), the dominant kernel's roofline figure
from HIP events recorded on the library's own stream, and the reference CPU assembly
timed on this host over bounded samples (one pinned core, median of 3, plus an
all-cores upper bound), all CPU timings taken before the GPU is first used.

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
J2_AMPLITUDE_40 = 1.2864e-4  # smooth-field amplitude at which ~40 % of the config-3 points yield


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


def kernel_sources_sha() -> str:
    """Hash of the library's kernel and launch sources (and its C header): traffic
    counted on other sources is not reported."""
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "calculixpp_amd", "csrc")
    files = sorted(os.path.join(csrc, f) for f in os.listdir(csrc)
                   if f.endswith((".hip", ".cpp", ".hpp")))
    for path in files + [os.path.join(ROOT, "include", "cxg.h")]:
        with open(path, "rb") as f:
            h.update(os.path.basename(path).encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def load_traffic(workload: str, variant: int):
    """HBM bytes per kernel launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    copy-calibrated, as committed in profiles/traffic.json: {kernel: bytes}, plus
    "step". None when not measured for this variant or when the record was counted on
    kernel sources other than these (its sources_sha16 differs)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("asm_variant") != variant:
        return None
    if rec.get("sources_sha16") != kernel_sources_sha():
        return None
    return rec


# Kernels of each counted workload: the library's timer name -> the kernel's name in
# rocprofv3's records (default options: asm_variant 7, gradient form for C3D20R).
_TRAFFIC_KERNELS = {
    "C3D20-64": {"elem_iso_up_C3D20": "k_elem_iso<20, 27, 3, 4, true, 1, true>",
                 "row_gather_up": "k_row_gather<1, false, false, true, 0, false>"},
    "C3D20R-64": {"elem_hform_C3D20R": "k_elem_hform<20, 8>",
                  "row_gather_hform": "k_row_gather<1, false, false, true, 8, false>"},
    "J2-C3D8-128x128x64": {"elem_j2_up_C3D8": "k_elem_j2<8, 8, 8, true>",
                           "row_gather_j2_up": "k_row_gather<0, false, false, true, 0, false>"},
}


def _short_kernel(name: str) -> str:
    """k_name<template args> of a rocprofv3 kernel name."""
    i = name.find("::k_")
    i = i + 2 if i >= 0 else 0
    depth, j = 0, i
    while j < len(name):
        if name[j] == "<":
            depth += 1
        elif name[j] == ">":
            depth -= 1
        elif name[j] == "(" and depth == 0:
            break
        j += 1
    return name[i:j].strip()


def _pmc_per_dispatch(folder: str, counter: str) -> dict:
    """{short kernel name: mean counter value per dispatch} of a rocprofv3 --pmc run."""
    import csv
    import glob
    sums, counts = {}, {}
    for path in glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                k = _short_kernel(r["Kernel_Name"])
                sums[k] = sums.get(k, 0.0) + float(r["Counter_Value"])
                counts[k] = counts.get(k, 0) + 1
    return {k: sums[k] / counts[k] for k in sums}


def ingest_traffic(root: str, copy_mib: int) -> dict:
    """profiles/traffic.json records from rocprofv3 --pmc runs under root: calib_fetch /
    calib_write (bench.py --copy-calibration copy_mib, one pass per counter) and, per
    workload W of _TRAFFIC_KERNELS, W_fetch / W_write. FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) are scaled by what they read for the plain copy (bytes copied / counted)."""
    copy_bytes = copy_mib * (1 << 20)
    cf = max(_pmc_per_dispatch(os.path.join(root, "calib_fetch"), "FETCH_SIZE").values())
    cw = max(_pmc_per_dispatch(os.path.join(root, "calib_write"), "WRITE_SIZE").values())
    f_fetch, f_write = copy_bytes / (cf * 1024), copy_bytes / (cw * 1024)
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        data = {}
    sha = kernel_sources_sha()
    for wl, names in _TRAFFIC_KERNELS.items():
        fd, wd = os.path.join(root, wl + "_fetch"), os.path.join(root, wl + "_write")
        if not (os.path.isdir(fd) and os.path.isdir(wd)):
            continue
        fetch = _pmc_per_dispatch(fd, "FETCH_SIZE")
        write = _pmc_per_dispatch(wd, "WRITE_SIZE")
        kern, detail = {}, {}
        for timer, rk in names.items():
            if rk not in fetch or rk not in write:
                raise SystemExit(f"{wl}: no counter record for {rk} (have {sorted(fetch)})")
            fb, wb = fetch[rk] * 1024 * f_fetch, write[rk] * 1024 * f_write
            kern[timer] = int(round(fb + wb))
            detail[timer] = {"kernel": rk, "fetch_bytes": int(round(fb)),
                             "write_bytes": int(round(wb)),
                             "fetch_kib_counted": round(fetch[rk], 1),
                             "write_kib_counted": round(write[rk], 1)}
        data[wl] = {"asm_variant": 7, "sources_sha16": sha, "kernels": kern,
                    "step": int(sum(kern.values())), "detail": detail,
                    "calibration": {"copy_bytes": copy_bytes, "fetch_factor": round(f_fetch, 4),
                                    "write_factor": round(f_write, 4)}}
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    return data


# ---- CPU baselines: run in forked children before the GPU is first touched ----------

_BATCH_RECS = None  # the batch workload's meshes, generated once, shared with the forks

def _cpu_model() -> str:
    """The host CPU as /proc/cpuinfo names it, with its family and model numbers and the
    logical CPUs seen and allowed (a virtual machine may report only a vendor line, e.g.
    "AMD EPYC", as its model name)."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break  # the first processor's block
                key, _, val = line.partition(":")
                fields.setdefault(key.strip(), val.strip())
    except OSError:
        pass
    import platform
    name = fields.get("model name") or platform.processor() or "unknown"
    ids = ", ".join(f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields)
    return (f"{name} ({ids + '; ' if ids else ''}{os.cpu_count()} logical CPUs, "
            f"{len(_allowed_cpus())} allowed)")


def _allowed_cpus() -> list:
    try:
        return sorted(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover - non-Linux
        return list(range(os.cpu_count() or 1))


def _cpu_job(job: dict) -> dict:
    """One timed CPU run of the reference core (or the oracle restatement when the
    reference build is absent), in the calling process."""
    from calculixpp_amd import _lib, meshgen
    from oracle.oracle import Oracle, aos_state, ref_available
    kind = "reference" if ref_available() else "port"
    o = Oracle(kind)
    what = job["what"]
    if what in ("assembly", "mass", "recover"):
        fam = _lib.TYPE_NAMES[job["family"]]
        m = meshgen.block(job["n"], fam, seed_geo=job.get("seed", 0), seed_E=job.get("seed", 0) + 1)
        u = meshgen.smooth_field(m.coords, 2)
        if what == "assembly":
            fn = lambda: o.time_linear_static(m, u)
        elif what == "recover":
            def fn():
                t0 = time.perf_counter()
                o.recover_fields(m, u)
                return time.perf_counter() - t0
        else:
            from calculixpp_amd.model import Model
            mm = Model(m.coords, m.blocks, m.E, m.nu, rho=7.85e-9)

            def fn():
                t0 = time.perf_counter()
                o.assemble_mass(mm)
                return time.perf_counter() - t0
        units = m.n_elem
    elif what == "j2":
        m = meshgen.j2_block(*job["shape"])
        h = aos_state(meshgen.seeded_history(m.n_points, 3))
        u = meshgen.smooth_field(m.coords, 7, job["amplitude"])
        fn = lambda: o.time_material_tangent(m, u, h)
        units = m.n_elem
    elif what == "batch":
        recs = _BATCH_RECS if _BATCH_RECS is not None else meshgen.batch_meshes(job["meshes"])

        def step(r):  # one mesh's step: seconds and element evaluations
            mdl = r["model"]
            if r["j2"]:
                return (sum(o.time_material_tangent(mdl, uu, aos_state(r["history"]))
                            for uu in r["u"]), len(r["u"]) * mdl.n_elem)
            return o.time_linear_static(mdl, r["u"][0]), mdl.n_elem

        sec, ev, chosen = 0.0, 0, []
        for r in recs:  # the batch's meshes in order, until about job["seconds"] of work,
            if sec >= job["seconds"]:  # leaving out meshes too large for the budget
                break
            if r["model"].n_elem > job["max_elements"]:
                continue
            t, n = step(r)
            sec, ev = sec + t, ev + n
            chosen.append(r)
        # the same sample again for the median of job["reps"] runs
        times = [sec] + [sum(step(r)[0] for r in chosen) for _ in range(job.get("reps", 1) - 1)]
        return {"kind": kind, "units": ev, "times": times, "meshes": len(chosen)}
    else:
        raise ValueError(what)
    if job.get("barrier") is not None:
        job["barrier"].wait()
    times = [fn() for _ in range(job.get("reps", 1))]
    return {"kind": kind, "units": units, "times": times}


def _pinned(job: dict, cpu: int, out):
    os.sched_setaffinity(0, {cpu})
    try:
        out.put(_cpu_job(job))
    except Exception as e:  # reported, never fatal to the GPU legs
        out.put({"error": f"{type(e).__name__}: {e}"})


def cpu_one_core(job: dict) -> dict:
    """job on one pinned host core (the first this process may use), in a fork."""
    import multiprocessing as mp
    ctx = mp.get_context("fork")
    q = ctx.Queue()
    p = ctx.Process(target=_pinned, args=(job, _allowed_cpus()[0], q))
    p.start()
    res = q.get()
    p.join()
    return res


def cpu_all_cores(job: dict, max_procs: int = 16) -> dict:
    """P = min(cores this process may use, max_procs) independent processes, one per
    core, each timing the job on its own slab (its own seeds) after a common barrier;
    throughput = total units / the slowest process's time (no merge: an upper bound)."""
    import multiprocessing as mp
    cpus = _allowed_cpus()[:max_procs]
    ctx = mp.get_context("fork")
    q = ctx.Queue()
    bar = ctx.Barrier(len(cpus))
    procs = []
    for k, c in enumerate(cpus):
        j = dict(job, seed=100 + 7 * k, barrier=bar, reps=1)
        procs.append(ctx.Process(target=_pinned, args=(j, c, q)))
    for p in procs:
        p.start()
    res = [q.get() for _ in procs]
    for p in procs:
        p.join()
    errs = [r["error"] for r in res if "error" in r]
    if errs:
        return {"error": errs[0]}
    units = sum(r["units"] for r in res)
    wall = max(r["times"][0] for r in res)
    return {"value": units / wall, "processes": len(cpus), "units": units, "seconds": wall,
            "kind": res[0]["kind"]}


def baseline_record(res: dict, unit: str, sample: str, all_cores=None) -> dict:
    """cpu_baseline object: one pinned core, median of the runs, plus the all-cores
    upper bound when measured."""
    import statistics
    if "error" in res:
        return {"value": None, "unit": unit, "cores": 1, "kind": None, "error": res["error"]}
    t = statistics.median(res["times"])
    out = {"value": res["units"] / t, "unit": unit, "cores": 1, "kind": res["kind"],
           "sample": f"{sample}; one pinned host core, median of {len(res['times'])} runs "
                     f"({', '.join(f'{x:.2f}' for x in res['times'])} s)",
           "cpu_model": _cpu_model()}
    if all_cores is not None:
        out["all_cores"] = dict(all_cores, unit=unit, cpu_model=out["cpu_model"],
                                note="P independent processes on disjoint samples, no merge: "
                                     "an upper bound")
    return out


def cpu_phase(args) -> dict:
    """Every CPU baseline of the run, before any GPU work."""
    out = {}
    if args.no_cpu:
        return out
    log = lambda m: print(m, file=sys.stderr, flush=True)
    n = args.cpu_n
    for fam in ("C3D20", "C3D20R"):
        job = {"what": "assembly", "family": fam, "n": n, "reps": 3}
        one = cpu_one_core(job)
        allc = cpu_all_cores(job) if not args.no_all_cores else None
        out[fam] = baseline_record(one, "elements/s",
                                   f"{n}^3 {fam} block of the same seeds ({n ** 3} elements, "
                                   "assemble_linear_static + internal_force)", allc)
        log(f"[bench] cpu {fam}: {out[fam]['value']} elements/s on one core")
    if not args.no_j2:
        shape = (32, 32, 16)
        one = cpu_one_core({"what": "j2", "shape": shape, "amplitude": J2_AMPLITUDE_40, "reps": 3})
        out["j2"] = baseline_record(one, "elements/s",
                                    f"{shape[0]}x{shape[1]}x{shape[2]} C3D8 J2 block, same seeds, "
                                    "amplitude of 40 % yield, assemble_material_tangent")
    if not args.no_recover:
        one = cpu_one_core({"what": "recover", "family": args.family, "n": n, "reps": 3})
        out["recover"] = baseline_record(one, "elements/s",
                                         f"{n}^3 {args.family} block, recover_fields")
    if not args.no_mass:
        one = cpu_one_core({"what": "mass", "family": args.family, "n": n, "reps": 3})
        out["mass"] = baseline_record(one, "elements/s", f"{n}^3 {args.family} block, assemble_mass")
    if not args.no_batch:
        global _BATCH_RECS
        if _BATCH_RECS is None:
            from calculixpp_amd import meshgen
            _BATCH_RECS = meshgen.batch_meshes(args.batch_meshes)
        one = cpu_one_core({"what": "batch", "meshes": args.batch_meshes, "seconds": 6.0,
                            "max_elements": 5000, "reps": 3})
        if "error" not in one:
            out["batch"] = baseline_record(one, "element evaluations/s",
                                           f"the batch's first {one['meshes']} meshes of at most "
                                           f"5,000 elements ({one['units']} element evaluations, "
                                           "linear and J2 steps as in the batch)")
    return out


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


def j2_newton(shape=(128, 128, 64), reps: int = 5, target: float = 0.40, cpu_rec=None) -> dict:
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
                       "traffic": (load_traffic("J2-C3D8-128x128x64",
                                                ctx.options.get("asm_variant", 7)) or {}).get("step")}
    if cpu_rec and cpu_rec.get("value"):
        out["cpu_baseline"] = cpu_rec
        out["gpu_over_cpu"] = out["elements_per_s"] / cpu_rec["value"]
    return out


def batch_roofline(step, ms: float) -> dict:
    """Algorithmic FLOP and bytes of one batch step, summed per family: every linear
    element once (K) plus its residual per load case, every J2 element per load case
    (K_t, f_int and 300 FLOP per Gauss point of return map); bytes as the assembly formula
    per mesh set (K values, coordinates, connectivity, E; u read and f_int written per
    case) plus the J2 history read and written (208 bytes per point per case)."""
    from calculixpp_amd.model import GAUSS, NODES
    flop = byt = 0.0
    nc = step.n_cases
    for kind, b, asm in (("lin", step.lin, getattr(step, "lin_asm", None)),
                         ("j2", step.j2, getattr(step, "j2_asm", None))):
        if b is None:
            continue
        m = b.model
        for t, c in m.blocks:
            nn, ng, ne = NODES[t], GAUSS[t], c.shape[0]
            k = flops_per_element(nn, ng, False) * ne
            r = (flops_per_element(nn, ng, True) - flops_per_element(nn, ng, False)) * ne
            if kind == "lin":
                flop += k + nc * r
                byt += (4 * nn + 8) * ne
            else:
                flop += nc * (k + r + J2_FLOP_PER_POINT * ng * ne)
                byt += nc * (4 * nn + 8) * ne
        reps = 1 if kind == "lin" else nc
        byt += reps * (8 * asm.pattern.nnz + 24 * m.n_nodes) + nc * 48 * m.n_nodes
        if kind == "j2":
            byt += nc * 208 * m.n_points
    t_f, t_b = flop / (PEAK_FP64 * 1e12), byt / (PEAK_HBM * 1e9)
    return {"bound": "mfma" if t_f >= t_b else "hbm", "flop_per_step": flop,
            "algorithmic_bytes_per_step": byt, "t_attainable_ms": round(max(t_f, t_b) * 1e3, 4),
            "frac": round(max(t_f, t_b) * 1e3 / ms, 4),
            "achieved_TFLOPs": round(flop / (ms * 1e-3) / 1e12, 2),
            "achieved_GBps": round(byt / (ms * 1e-3) / 1e9, 1)}


def batch_step(recs, reps: int = 3, cpu_rec=None, log=None) -> dict:
    """The batched workload: 512 mixed C3D8/C3D20/C3D6 boxes with log-uniform element
    counts in [100, 50000] (seed 4), half of them J2, 4 load cases each. One step = the
    linear meshes' K plus f_int for the 4 cases, and for the J2 meshes per case the return
    map, K_t and f_int; the elements of every mesh go through one launch per family
    (asm_variant 7) or one persistent launch for all families (asm_variant 5), whichever
    autotune.tune_batch measures faster. Throughput counts element evaluations (a linear
    element once, a J2 element once per case)."""
    import torch
    from calculixpp_amd import autotune
    from calculixpp_amd.batch import BatchStep
    from calculixpp_amd import fem
    t1 = time.perf_counter()
    ctx = fem.context()
    step = BatchStep(recs, ctx, variant=7)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    if log:
        log(f"[bench] batch: {len(recs)} meshes uploaded + patterns in {t2 - t1:.1f}s")
    tune = autotune.tune_batch(step.run, ctx, log=log)
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
    best = min(eager_ms, graph_ms)
    out = {"workload": f"{len(recs)} mixed C3D8/C3D20/C3D6 meshes, log-uniform 100..50000 "
                       f"elements, half J2, {step.n_cases} load cases",
           "meshes": len(recs), "elements": n_lin + n_j2, "linear_elements": n_lin,
           "j2_elements": n_j2, "element_evaluations": evals, "nnz": nnz,
           "ms_per_step": eager_ms, "graph_ms_per_step": graph_ms,
           "evaluations_per_s": evals / (best * 1e-3),
           "roofline": batch_roofline(step, best),
           "autotune": tune, "kernels_one_graph_plus_one_eager_step": kern}
    ctx.set_option("asm_variant", 7)
    del step
    torch.cuda.empty_cache()
    if cpu_rec and cpu_rec.get("value"):
        out["cpu_baseline"] = cpu_rec
        out["gpu_over_cpu"] = out["evaluations_per_s"] / cpu_rec["value"]
    return out


def recover_leg(model, u, family: int, reps: int = 5, cpu_rec=None) -> dict:
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
    if cpu_rec and cpu_rec.get("value"):
        out["cpu_baseline"] = cpu_rec
        out["gpu_over_cpu"] = out["elements_per_s"] / cpu_rec["value"]
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


def mass_geometric_leg(asm, family: int, reps: int = 3, cpu_rec=None) -> dict:
    """Consistent mass, lumped mass and geometric stiffness of the assembly workload's
    mesh (density 7.85e-9, seeded Gauss-point stress), from HIP events on the library's
    stream. Both matrices are k_ab I3 blocks, written through the element arena and the
    row gather straight onto the same-direction pattern (row (a, i): columns (b, i)), the
    reference's key set when every entry is live; the structural-pattern gather (all
    nnz values, two thirds zeros) and its compaction are timed beside it, by wall clock
    (the compaction synchronizes). Algorithmic
    bytes: the values and live flags written (9 per entry) plus coordinates,
    connectivity and one value per element or point."""
    import statistics
    import numpy as np
    import torch
    from calculixpp_amd.model import GAUSS, NODES
    dev = asm.ctx.device
    model = asm.model
    rho = torch.full((model.n_elem,), 7.85e-9, dtype=torch.float64, device=dev)
    gs = torch.from_numpy(np.random.default_rng(45).normal(scale=100.0, size=(model.n_points, 6))).to(dev)
    ctx = asm.ctx
    nnz = asm.pattern.nnz
    nnz_sd = int(asm.same_direction()[1].numel())
    vs = torch.empty(max(nnz_sd, 1), dtype=torch.float64, device=dev)
    ls = torch.empty(max(nnz_sd, 1), dtype=torch.uint8, device=dev)
    vals, lv = asm.new_values(), asm.new_live()

    def structural():  # the path the same-direction rows replace: all nnz, then compaction
        asm.assemble_mass(vals, False, rho, live=lv)
        asm.compact(vals, lv)

    runs = {"mass": lambda: asm.assemble_mass_sd(vs, ls, False, rho),
            "mass_lumped": lambda: asm.assemble_mass_sd(vs, ls, True, rho),
            "geometric": lambda: asm.assemble_geometric_sd(vs, ls, gs),
            "mass_structural": structural}
    nn, ngp = NODES[family], GAUSS[family]
    out = {"workload": "mass (consistent, lumped) and geometric stiffness of the assembly "
                       "block onto its same-direction pattern (structural pattern beside it)",
           "elements": model.n_elem, "nnz_same_direction": nnz_sd, "nnz_structural": nnz}
    for key, fn in runs.items():
        fn()
        times, walls = [], []
        for _ in range(reps):
            torch.cuda.synchronize()
            ctx.profile_reset()
            ctx.profile(True)
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            walls.append((time.perf_counter() - t0) * 1e3)
            ctx.profile(False)
            times.append(sum(v[1] for v in ctx.profile_stats().values()))
        # the timed kernels; the structural path adds its compaction (untimed kernels, a
        # host synchronization), so its figure is the wall clock
        ms = statistics.median(walls) if key == "mass_structural" else statistics.median(times)
        extra = 48 * ngp * model.n_elem if key == "geometric" else 8 * model.n_elem
        n_out = nnz if key == "mass_structural" else nnz_sd
        per = 8 if key == "mass_structural" else 9
        byt = per * n_out + 24 * model.n_nodes + 4 * nn * model.n_elem + extra
        t_att = byt / (PEAK_HBM * 1e9)
        out[key] = {"ms": ms, "wall_ms": round(statistics.median(walls), 4),
                    "elements_per_s": model.n_elem / (ms * 1e-3),
                    "kernels_ms": {k: round(v[1], 4) for k, v in ctx.profile_stats().items()},
                    "roofline": {"bound": "hbm", "algorithmic_bytes": byt,
                                 "t_attainable_ms": round(t_att * 1e3, 4),
                                 "frac": round(t_att / (ms * 1e-3), 4)}}
        if key != "mass_structural":  # the key set is the whole pattern when all are live
            out[key]["all_live"] = bool(ls[:nnz_sd].all()) if nnz_sd else True
    del vals, lv, vs, ls
    torch.cuda.synchronize()
    if cpu_rec and cpu_rec.get("value"):
        out["cpu_baseline"] = cpu_rec
        out["gpu_over_cpu"] = out["mass"]["elements_per_s"] / cpu_rec["value"]
    return out


def assembly_leg(family_name: str, n: int, residual: bool, steps: int, warmup: int,
                 variant: int = 0, patch: int = 0, options=(), log=None) -> dict:
    """One assembly workload: the n^3 block of family_name (seeds 0, 1, 2), K values
    (+ internal force), inputs resident in HBM. Wall-clock throughput over `steps`
    synchronized steps, then per-kernel durations from HIP events on the library's
    stream. Roofline figures: per kernel against its own bound (the element kernel's
    algorithmic FP64 count over its time, the row gather's algorithmic bytes -- K values
    written, u read and f_int written -- over its time), the dominant (longest) kernel
    being the headline, and over the step as max(F/peak_fp64, B/peak_hbm) / t."""
    import torch
    from calculixpp_amd import _lib, autotune, fem, meshgen
    from calculixpp_amd.model import GAUSS, NODES
    family = _lib.TYPE_NAMES[family_name]
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    model = meshgen.block(n, family)
    u = torch.from_numpy(meshgen.smooth_field(model.coords, 2)).to(dev) if residual else None
    t1 = time.perf_counter()
    if variant and patch:
        choice = {"asm_variant": variant, "patch_size": patch, "trials": []}
    else:
        choice = autotune.tune_linear(model, u, cache_path=None, log=log)
    t2 = time.perf_counter()
    for opt in options:
        key, val = opt.split("=", 1)
        fem.context().set_option(key, int(val))
    asm = autotune.make_assembler(model, choice["asm_variant"], choice["patch_size"])
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    if log:
        log(f"[bench] {family_name}: mesh {t1 - t0:.1f}s, tune {t2 - t1:.1f}s, "
            f"upload+colour+pattern {t3 - t2:.1f}s: {model.n_elem} elements, {model.n_nodes} "
            f"nodes, nnz {asm.pattern.nnz}, variant {choice['asm_variant']}")
    vals = asm.new_values()
    f_int = torch.empty(3 * model.n_nodes, dtype=torch.float64, device=dev) if residual else None
    asm.assemble(vals, u, f_int)  # set-up: arenas allocated, gather order built
    for _ in range(warmup):
        asm.assemble(vals, u, f_int)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(steps):
        asm.assemble(vals, u, f_int)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - ts) * 1e3 / steps
    value = model.n_elem / (ms_step * 1e-3)
    if log:
        log(f"[bench] {family_name}: {ms_step:.3f} ms/step, {value:.4g} elements/s")
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
    # the staged assembly (asm_variant 8) runs its element and gather launches on two
    # overlapping streams: its step is the whole-call timer, each kernel's time the sum
    # of its launches in one step
    staged = stats.pop("staged_assembly", None)
    step_ms = staged[1] / reps if staged else sum(v[1] for v in stats.values()) / reps
    nn, ngp = NODES[family], GAUSS[family]
    fpe = flops_per_element(nn, ngp, residual)
    flop = fpe * model.n_elem
    bpa = bytes_per_assembly(asm.pattern.nnz, model.n_nodes, model.n_elem, nn, residual)
    gather_bytes = 8 * asm.pattern.nnz + (48 * model.n_nodes if residual else 0)
    t_flop, t_byte = flop / (PEAK_FP64 * 1e12), bpa / (PEAK_HBM * 1e9)
    traffic = load_traffic(f"{family_name}-{n}", choice["asm_variant"])
    per_kernel = {}
    for k, v in stats.items():
        avg = v[1] / reps * 1e-3  # seconds per step
        if k.startswith("row_gather"):
            per_kernel[k] = {"bound": "hbm", "algorithmic_bytes": gather_bytes,
                             "achieved": round(gather_bytes / avg / 1e9, 1), "peak": PEAK_HBM,
                             "unit": "GB/s", "frac": round(gather_bytes / avg / 1e9 / PEAK_HBM, 4)}
        elif k.startswith("elem_hform"):
            # gradient form: the slot (lambda, mu, 3n gradients per point) written, the
            # connectivity and (u given) the per-element displacement copy
            hb = model.n_elem * (8 * (2 + 3 * nn * ngp) + 4 * nn + (48 * nn if residual else 0))
            per_kernel[k] = {"bound": "hbm", "algorithmic_bytes": hb,
                             "achieved": round(hb / avg / 1e9, 1), "peak": PEAK_HBM,
                             "unit": "GB/s", "frac": round(hb / avg / 1e9 / PEAK_HBM, 4)}
        else:
            per_kernel[k] = {"bound": "mfma", "algorithmic_flop": flop,
                             "achieved": round(flop / avg / 1e12, 2), "peak": PEAK_FP64,
                             "unit": "TFLOP/s", "frac": round(flop / avg / 1e12 / PEAK_FP64, 4)}
        per_kernel[k]["ms_per_step"] = round(v[1] / reps, 4)
        per_kernel[k]["launches_per_step"] = v[0] // reps
        per_kernel[k]["traffic"] = traffic["kernels"].get(k) if traffic else None
    dominant = max(stats, key=lambda k: stats[k][1])
    head = dict(per_kernel[dominant])
    head.update({"kernel": dominant,
                 "traffic_note": ("HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE + "
                                  "WRITE_SIZE (copy-calibrated), profiles/traffic.json, counted "
                                  "on these kernel sources") if traffic else
                 "no counter record for these kernel sources",
                 "kernels": per_kernel,
                 "step": {"frac": round(max(t_flop, t_byte) / (step_ms * 1e-3), 4),
                          "staged": staged is not None,
                          "bound": "mfma" if t_flop >= t_byte else "hbm",
                          "step_kernel_ms": round(step_ms, 4),
                          "t_attainable_ms": round(max(t_flop, t_byte) * 1e3, 4),
                          "flop_per_element": fpe, "flop_per_step": flop,
                          "algorithmic_bytes_per_step": bpa,
                          "achieved_TFLOPs": round(flop / (step_ms * 1e-3) / 1e12, 2),
                          "hbm_algorithmic_GBps": round(bpa / (step_ms * 1e-3) / 1e9, 1),
                          "traffic": traffic.get("step") if traffic else None}})
    return {"model": model, "u": u, "asm": asm, "family": family, "value": value,
            "ms_per_step": ms_step, "choice": choice, "kernels": kernels, "roofline": head,
            "config": {"workload": f"{n}^3 {family_name} jittered block: CSR K values"
                                   f"{' + internal force' if residual else ''}",
                       "elements": model.n_elem, "nodes": model.n_nodes,
                       "dof": 3 * model.n_nodes, "nnz": asm.pattern.nnz,
                       "colours": asm.mesh.n_colors, "asm_variant": choice["asm_variant"],
                       "patch_size": choice["patch_size"]}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--family", default="C3D20")
    ap.add_argument("--cpu-n", type=int, default=10, help="edge of the CPU-baseline samples")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-all-cores", action="store_true",
                    help="skip the all-cores upper bound of the CPU baseline")
    ap.add_argument("--no-residual", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adj-reps", type=int, default=20)
    ap.add_argument("--no-second", action="store_true",
                    help="skip the second assembly family (C3D20R, or C3D20 for --family C3D20R)")
    ap.add_argument("--no-j2", action="store_true", help="skip the J2 Newton-iteration leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-meshes leg")
    ap.add_argument("--no-recover", action="store_true", help="skip the field-recovery leg")
    ap.add_argument("--no-mass", action="store_true",
                    help="skip the mass / geometric-stiffness leg")
    ap.add_argument("--no-pcg", action="store_true", help="skip the device-solve leg")
    ap.add_argument("--batch-meshes", type=int, default=512)
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="extra cxg_set_option for the library context (repeatable)")
    ap.add_argument("--only", choices=["j2", "batch", "adjoint", "pcg", "assembly"], default=None,
                    help="run one leg alone and print its JSON (profiling)")
    ap.add_argument("--variant", type=int, default=0, help="force asm_variant (0 = autotune)")
    ap.add_argument("--patch", type=int, default=0, help="force patch_size (0 = autotune)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    ap.add_argument("--ingest-traffic", default=None, metavar="DIR",
                    help="write profiles/traffic.json from rocprofv3 --pmc runs under DIR "
                         "(no GPU use; see ingest_traffic)")
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if args.ingest_traffic:
        data = ingest_traffic(args.ingest_traffic, args.copy_calibration or 1024)
        print(json.dumps({k: v.get("step") for k, v in data.items()}), flush=True)
        return 0
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
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding of this path is a design, not a build", file=sys.stderr)
        return 2
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
            res = j2_newton(reps=args.steps)
        elif args.only == "pcg":
            res = pcg_leg(args.n)
        elif args.only == "batch":
            from calculixpp_amd import meshgen
            res = batch_step(meshgen.batch_meshes(args.batch_meshes), reps=args.steps, log=log)
        elif args.only == "assembly":
            leg = assembly_leg(args.family, args.n, not args.no_residual, args.steps, args.warmup,
                               args.variant, args.patch or (8 if args.variant else 0), args.option,
                               log)
            res = {k: leg[k] for k in ("value", "ms_per_step", "kernels", "roofline", "config")}
        else:
            res = adjoint_ratio(args.adj_reps)
        print(json.dumps(res), flush=True)
        return 0

    # CPU baselines first, in forked children, before anything touches the GPU
    cpu = cpu_phase(args)
    import torch
    torch.cuda.set_device(0)
    residual = not args.no_residual
    leg = assembly_leg(args.family, args.n, residual, args.steps, args.warmup, args.variant,
                       args.patch, args.option, log)
    model, u, asm, family = leg["model"], leg["u"], leg["asm"], leg["family"]
    head = {k: leg[k] for k in ("value", "ms_per_step", "config", "choice", "roofline", "kernels")}
    main_cpu = cpu.get(args.family)

    second = None
    if not args.no_second:
        fam2 = "C3D20R" if args.family != "C3D20R" else "C3D20"
        # the first leg's device buffers stay for the recovery and mass legs below
        leg2 = assembly_leg(fam2, args.n, residual, args.steps, args.warmup, args.variant,
                            args.patch, args.option, log)
        second = {"family": fam2, "value": leg2["value"], "unit": "elements/s",
                  "ms_per_step": leg2["ms_per_step"], "config": leg2["config"],
                  "roofline": leg2["roofline"], "kernels": leg2["kernels"],
                  "cpu_baseline": cpu.get(fam2)}
        if second["cpu_baseline"] and second["cpu_baseline"].get("value"):
            second["gpu_over_cpu"] = leg2["value"] / second["cpu_baseline"]["value"]
        del leg2
        torch.cuda.empty_cache()

    adjoint = None if args.no_adjoint else adjoint_ratio(args.adj_reps)
    if adjoint:
        log(f"[bench] adjoint/forward {adjoint['ratio']:.3f} ({adjoint['vjp_ms']:.4f} / "
            f"{adjoint['forward_ms']:.4f} ms)")
    j2 = None if args.no_j2 else j2_newton(cpu_rec=cpu.get("j2"))
    if j2:
        log(f"[bench] J2 Newton iteration {j2['ms_per_iteration']:.3f} ms "
            f"({j2['elements_per_s']:.4g} elements/s, yield {j2['yield_fraction']:.3f})")
    rec = None
    if not args.no_recover and u is not None:
        rec = recover_leg(model, u.reshape(-1, 3), family, cpu_rec=cpu.get("recover"))
        log(f"[bench] recover_fields {rec['ms']:.3f} ms ({rec['elements_per_s']:.4g} elements/s)")
    mass = None if args.no_mass else mass_geometric_leg(asm, family, cpu_rec=cpu.get("mass"))
    if mass:
        log(f"[bench] mass {mass['mass']['ms']:.3f} ms, lumped {mass['mass_lumped']['ms']:.3f} "
            f"ms, geometric {mass['geometric']['ms']:.3f} ms")
    del asm, leg
    torch.cuda.empty_cache()
    pcg = None if args.no_pcg else pcg_leg(args.n)
    if pcg:
        log(f"[bench] pcg {pcg['ms_per_iteration']:.3f} ms per iteration "
            f"({pcg['roofline']['achieved_GBps']:.0f} GB/s)")
    batch = None
    if not args.no_batch:
        from calculixpp_amd import meshgen
        recs = _BATCH_RECS if _BATCH_RECS is not None else meshgen.batch_meshes(args.batch_meshes)
        batch = batch_step(recs, cpu_rec=cpu.get("batch"), log=log)
        log(f"[bench] batch step {batch['ms_per_step']:.2f} ms eager, "
            f"{batch['graph_ms_per_step']:.2f} ms graph ({batch['evaluations_per_s']:.4g} "
            "element evaluations/s)")
    out = {
        "metric": "C3D20 K-assembly elements/s + % FP64/HBM roofline; adjoint/forward cost",
        "value": head["value"], "unit": "elements/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": head["config"],
        "autotune": head["choice"]["trials"],
        "roofline": head["roofline"],
        "cpu_baseline": main_cpu,
        "kernels": head["kernels"],
        "second_family": second,
        "adjoint": adjoint,
        "j2_newton": j2,
        "recover": rec,
        "mass_geometric": mass,
        "pcg": pcg,
        "batch": batch,
        "gpu_over_cpu": (head["value"] / main_cpu["value"]) if main_cpu and main_cpu.get("value")
        else None,
        "kernel_sources_sha16": kernel_sources_sha(),
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
