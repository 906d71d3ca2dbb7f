"""Autotuner of the assembly launch configuration.

Two knobs decide how the stiffness assembly maps onto the GPU:
  asm_variant  1 = per-pair scatter (one element per wave, all Gauss-point
                   gradients in LDS), 2/3 = M-identity kernel with row-coalesced
                   scatter at 4/3 waves per SIMD, 4 = two-phase (element matrices
                   to an arena, rows gathered by their node), 6 = two-phase with
                   M = H^T H on the FP64 matrix cores, 7 = two-phase with only the
                   upper node-pair blocks in the arena (half the bytes), gathered
                   in a locality order of the nodes so that both readers of a
                   block find it in L2;
  patch_size   consecutive elements one workgroup walks (patch colouring keeps
               patches of one launch node-disjoint; longer patches reuse the CSR
               rows of shared nodes while they are still in cache, shorter ones
               give more parallelism).
The tuner times each candidate on the actual mesh (HIP events on the library's
stream) and caches the winner in a JSON file keyed by element family, element
count and nonzero count, so later runs on the same workload reuse it.

A batch of uneven meshes (batch.BatchStep) runs either as per-family launches with
upper blocks (asm_variant 7: the batch concatenates each family over its meshes, so
one family is one launch) or as one persistent launch (asm_variant 5) with two more
knobs:
  chunk_elems       elements per work item (the unit a workgroup takes from the
                    queue: small items balance better, large ones cost fewer atomics)
  persistent_waves  resident waves per compute unit of the persistent grid.
tune_batch times variant 7 and every persistent candidate on the actual batch step
and keeps the fastest; tune_persistent sweeps the persistent knobs alone.
"""
from __future__ import annotations

import json
import os
import time
from typing import Iterable, Optional

import torch

from . import fem
from .model import Model

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "calculixpp_amd", "autotune.json")
CANDIDATES = [(7, 8), (4, 8), (6, 8), (2, 8), (1, 8)]


def _key(model: Model) -> str:
    fams = "+".join(str(t) for t, _ in model.blocks)
    return f"linear:{fams}:{model.n_elem}:{model.n_nodes}"


def load_cache(path: str = DEFAULT_CACHE) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(cache: dict, path: str = DEFAULT_CACHE) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1)
    os.replace(tmp, path)


def make_assembler(model: Model, variant: int, patch: int,
                   ctx: Optional[fem.Context] = None) -> fem.Assembler:
    ctx = ctx or fem.context()
    ctx.set_option("asm_variant", variant)
    ctx.set_option("patch_size", patch)
    return fem.Assembler(model, ctx)


def time_assembler(asm: fem.Assembler, vals, u=None, f_int=None, steps: int = 3) -> float:
    """Mean milliseconds per assembly (after one warm-up)."""
    asm.assemble(vals, u, f_int)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        asm.assemble(vals, u, f_int)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def tune_linear(model: Model, u=None, candidates: Iterable = CANDIDATES, steps: int = 3,
                cache_path: Optional[str] = DEFAULT_CACHE, force: bool = False, log=None) -> dict:
    """Pick (asm_variant, patch_size) for this model; returns
    {"asm_variant", "patch_size", "ms", "trials": [...]}."""
    key = _key(model)
    cache = load_cache(cache_path) if cache_path else {}
    if not force and key in cache:
        return cache[key]
    ctx = fem.context()
    trials = []
    best = None
    for variant, patch in candidates:
        asm = make_assembler(model, variant, patch, ctx)
        vals = asm.new_values()
        f_int = None
        if u is not None:
            f_int = torch.empty(3 * model.n_nodes, dtype=torch.float64, device=ctx.device)
        ms = time_assembler(asm, vals, u, f_int, steps)
        trials.append({"asm_variant": variant, "patch_size": patch, "ms": ms})
        if log:
            log(f"[autotune] variant {variant} patch {patch}: {ms:.3f} ms")
        if best is None or ms < best["ms"]:
            best = {"asm_variant": variant, "patch_size": patch, "ms": ms}
        del asm, vals, f_int
        torch.cuda.empty_cache()
    best["trials"] = trials
    if cache_path:
        cache[key] = best
        save_cache(cache, cache_path)
    return best


PERSISTENT_CANDIDATES = [(1, 8), (2, 8), (4, 8), (8, 8), (16, 8), (4, 4), (4, 16), (8, 16)]


def tune_persistent(run, ctx: fem.Context, candidates: Iterable = PERSISTENT_CANDIDATES,
                    steps: int = 2, log=None) -> dict:
    """Pick (chunk_elems, persistent_waves) for a callable `run` that enqueues one step
    on ctx's stream with asm_variant 5 (e.g. batch.BatchStep.run); the winner is left
    set on ctx. Returns {"chunk_elems", "persistent_waves", "ms", "trials"}."""
    trials, best = [], None
    for chunk, waves in candidates:
        ctx.set_option("chunk_elems", chunk)
        ctx.set_option("persistent_waves", waves)
        run()  # rebuilds the work queue for this item size
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            run()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        trials.append({"chunk_elems": chunk, "persistent_waves": waves, "ms": ms})
        if log:
            log(f"[autotune] chunk {chunk} waves {waves}: {ms:.3f} ms")
        if best is None or ms < best["ms"]:
            best = {"chunk_elems": chunk, "persistent_waves": waves, "ms": ms}
    ctx.set_option("chunk_elems", best["chunk_elems"])
    ctx.set_option("persistent_waves", best["persistent_waves"])
    best["trials"] = trials
    return best


def tune_batch(run, ctx: fem.Context, candidates: Iterable = PERSISTENT_CANDIDATES,
               steps: int = 2, log=None) -> dict:
    """Pick how a batch step runs (`run` enqueues one step on ctx's stream, e.g.
    batch.BatchStep.run): per-family launches (asm_variant 7) or the persistent launch
    (asm_variant 5) at its best (chunk_elems, persistent_waves). The winner is left set
    on ctx. Returns {"asm_variant", "ms", "trials"} plus chunk_elems and
    persistent_waves for variant 5."""
    ctx.set_option("asm_variant", 7)
    run()  # arenas, gather order and descriptors for this variant
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    ms7 = (time.perf_counter() - t0) * 1e3 / steps
    if log:
        log(f"[autotune] per-family launches (variant 7): {ms7:.3f} ms")
    ctx.set_option("asm_variant", 5)
    pers = tune_persistent(run, ctx, candidates, steps, log)
    trials = [{"asm_variant": 7, "ms": ms7}] + [dict(asm_variant=5, **t) for t in pers["trials"]]
    if ms7 <= pers["ms"]:
        ctx.set_option("asm_variant", 7)
        best = {"asm_variant": 7, "ms": ms7}
    else:
        best = {"asm_variant": 5, "chunk_elems": pers["chunk_elems"],
                "persistent_waves": pers["persistent_waves"], "ms": pers["ms"]}
    best["trials"] = trials
    return best
