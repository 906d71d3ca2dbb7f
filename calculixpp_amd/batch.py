"""Many independent meshes of uneven size and mixed families assembled as one.

The reference assembles one Model per call (fem::assemble_linear_static,
src/fem/assembly.cpp:225-257; fem::assemble_material_tangent, :385-423). A batch of
independent meshes is concatenated here into one block-diagonal model: node ranges
are disjoint, each family's elements of every mesh form one block, so one element
launch per family (asm_variant 7, upper node-pair blocks; the gather order goes mesh by
mesh, each mesh's nodes in its own locality order) or one persistent launch for all
families (asm_variant 5) covers every element of every mesh, and one gather launch every
row (autotune.tune_batch picks between them by measurement). Each mesh's blocks keep
their order inside the family runs, so every node still
sums its incident elements in its own mesh's element order (the order of the
reference's duplicate summation) when the mesh lists its blocks in family order
C3D8, C3D20, C3D20R, C3D6 (meshgen.mixed_box does). Per-mesh results are views:
contiguous equation, nonzero and node ranges of the batch's arrays.

    b = Batch([m0, m1, ...])                # concatenated model + per-mesh maps
    rp, ci, v = b.csr(i, row_ptr, col_idx, vals)
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional

import numpy as np
import torch

from . import _lib, fem
from .model import GAUSS, Model

FAMILY_ORDER = (_lib.C3D8, _lib.C3D20, _lib.C3D20R, _lib.C3D6)


class Batch:
    """A list of models as one block-diagonal model (see the module docstring)."""

    def __init__(self, models: List[Model]):
        if not models:
            raise ValueError("empty batch")
        self.models = models
        n_nodes = np.array([m.n_nodes for m in models], np.int64)
        self.node_off = np.concatenate([[0], np.cumsum(n_nodes)])
        mat_off = np.concatenate([[0], np.cumsum([len(m.j2) for m in models])])
        any_j2 = any(m.j2_id is not None for m in models)
        blocks, E, nu, jid = [], [], [], []
        # super element / Gauss point index of every mesh element / Gauss point
        self.elem_map = [np.empty(m.n_elem, np.int64) for m in models]
        self.gp_map = [np.empty(m.n_points, np.int64) for m in models]
        se = sp = 0
        for fam in FAMILY_ORDER:
            conns = []
            for i, m in enumerate(models):
                le = lp = 0
                for t, c in m.blocks:
                    ng = GAUSS[t]
                    if t == fam:
                        n = c.shape[0]
                        conns.append(c + np.int32(self.node_off[i]))
                        self.elem_map[i][le: le + n] = se + np.arange(n)
                        self.gp_map[i][lp: lp + n * ng] = sp + np.arange(n * ng)
                        E.append(m.E[le: le + n])
                        nu.append(m.nu[le: le + n])
                        if any_j2:
                            if m.j2_id is None:
                                jid.append(np.full(n, -1, np.int32))
                            else:
                                j = m.j2_id[le: le + n]
                                jid.append(np.where(j >= 0, j + mat_off[i], -1).astype(np.int32))
                        se += n
                        sp += n * ng
                    le += c.shape[0]
                    lp += c.shape[0] * ng
            if conns:
                blocks.append((fam, np.concatenate(conns)))
        coords = np.concatenate([m.coords for m in models])
        j2 = [p for m in models for p in m.j2]
        for m in models:
            if len(m.spc_node) or len(m.cload_node) or (m.active is not None and not np.all(m.active)):
                raise NotImplementedError("batches carry no boundary conditions, loads or removals")
        self.model = Model(coords, blocks, np.concatenate(E), np.concatenate(nu),
                           j2_id=np.concatenate(jid) if any_j2 else None, j2=j2)
        # every DOF is free: mesh i's equations are 3*node_off[i] .. 3*node_off[i+1]
        self.eq_off = 3 * self.node_off

    @property
    def n_meshes(self) -> int:
        return len(self.models)

    def csr(self, i: int, row_ptr, col_idx, vals):
        """Mesh i's CSR (numpy row_ptr, col_idx, vals) out of the batch's arrays."""
        rp = _np(row_ptr)
        e0, e1 = int(self.eq_off[i]), int(self.eq_off[i + 1])
        a, b = int(rp[e0]), int(rp[e1])
        return rp[e0: e1 + 1] - a, _np(col_idx[a:b]) - e0, _np(vals[a:b])

    def nodal(self, i: int, f):
        """Mesh i's part of a full-DOF vector (3*n_nodes of the batch)."""
        return _np(f[int(self.eq_off[i]): int(self.eq_off[i + 1])])

    def stack_nodal(self, per_mesh) -> np.ndarray:
        """Concatenate per-mesh (n_nodes_i, 3) arrays into the batch's node order."""
        return np.concatenate([np.asarray(a, np.float64).reshape(-1, 3) for a in per_mesh])

    def stack_points(self, per_mesh) -> np.ndarray:
        """(13, P) batch history from per-mesh (13, P_i) histories (elastic meshes: None)."""
        out = np.zeros((_lib.STATE_COMPONENTS, self.model.n_points))
        for i, h in enumerate(per_mesh):
            if h is not None:
                out[:, self.gp_map[i]] = h
        return out

    def points(self, i: int, state) -> np.ndarray:
        """Mesh i's (13, P_i) part of a batch history."""
        return _np(state)[:, self.gp_map[i]]


def _np(a):
    return a.cpu().numpy() if torch.is_tensor(a) else np.asarray(a)


class BatchStep:
    """One step of a batch of meshes and load cases: the linear meshes' K with the
    internal force of every load case (one pass), and for the J2 meshes, per load case,
    the return map, K_t, f_int and trial history, all from the same committed state.

    records: dicts with "model", "j2" (bool), "u" (n_cases, n_nodes, 3) and, for J2
    meshes, "history" (13, n_points) — the shape meshgen.batch_meshes returns."""

    def __init__(self, records, ctx: Optional[fem.Context] = None, variant: int = 7):
        self.ctx = ctx or fem.context()
        self.ctx.set_option("asm_variant", variant)
        dev = self.ctx.device
        self.n_cases = int(records[0]["u"].shape[0])
        lin = [r for r in records if not r["j2"]]
        j2 = [r for r in records if r["j2"]]
        self.lin_idx = [i for i, r in enumerate(records) if not r["j2"]]
        self.j2_idx = [i for i, r in enumerate(records) if r["j2"]]
        self.lin = self.j2 = None
        if lin:
            self.lin = Batch([r["model"] for r in lin])
            self.lin_asm = fem.Assembler(self.lin.model, self.ctx)
            n3 = 3 * self.lin.model.n_nodes
            U = np.stack([self.lin.stack_nodal([r["u"][c] for r in lin]).ravel()
                          for c in range(self.n_cases)])
            self.lin_u = torch.from_numpy(U).to(dev)
            self.lin_vals = self.lin_asm.new_values()
            self.lin_f = torch.empty((self.n_cases, n3), dtype=torch.float64, device=dev)
        if j2:
            self.j2 = Batch([r["model"] for r in j2])
            self.j2_asm = fem.TangentAssembler(self.j2.model, self.ctx)
            n3 = 3 * self.j2.model.n_nodes
            self.j2_u = [torch.from_numpy(self.j2.stack_nodal([r["u"][c] for r in j2]).ravel().copy()).to(dev)
                         for c in range(self.n_cases)]
            committed = torch.from_numpy(self.j2.stack_points([r["history"] for r in j2])).to(dev)
            self.j2_mp = []
            for c in range(self.n_cases):
                mp = fem.MaterialPoints(self.j2.model.n_points, dev, committed=None)
                mp.committed = committed  # shared, read-only during the step
                self.j2_mp.append(mp)
            self.j2_vals = [self.j2_asm.new_values() for _ in range(self.n_cases)]
            self.j2_f = [torch.empty(n3, dtype=torch.float64, device=dev) for _ in range(self.n_cases)]
        self.graphs = []

    @property
    def elements(self) -> int:
        return sum(b.model.n_elem for b in (self.lin, self.j2) if b is not None)

    def run(self) -> None:
        """The step through the eager entry points."""
        if self.lin is not None:
            a = self.lin_asm
            self.ctx.check(self.ctx.lib.cxg_assemble_linear_cases(
                self.ctx.h, a.pattern.h, fem._ptr(a.mesh.coords), fem._ptr(a.mesh.E),
                fem._ptr(a.mesh.nu), fem._ptr(self.lin_vals), self.n_cases, fem._ptr(self.lin_u),
                fem._ptr(self.lin_f), None))
        if self.j2 is not None:
            for c in range(self.n_cases):
                self.j2_asm.assemble_tangent(self.j2_u[c], self.j2_mp[c], self.j2_vals[c], self.j2_f[c])

    def capture(self) -> None:
        """Capture the step as HIP graphs (one for the linear meshes, one per J2 load
        case); launch() then replays them."""
        self.run()  # prepares arenas, work queues and the material table
        self.graphs = []
        if self.lin is not None:
            a = self.lin_asm
            h = C.c_void_p()
            self.ctx.check(self.ctx.lib.cxg_graph_capture_linear_cases(
                self.ctx.h, a.pattern.h, fem._ptr(a.mesh.coords), fem._ptr(a.mesh.E),
                fem._ptr(a.mesh.nu), fem._ptr(self.lin_vals), self.n_cases, fem._ptr(self.lin_u),
                fem._ptr(self.lin_f), C.byref(h)))
            self.graphs.append(fem.Graph(self.ctx, h, (a, self.lin_vals, self.lin_u, self.lin_f)))
        if self.j2 is not None:
            for c in range(self.n_cases):
                self.graphs.append(self.j2_asm.capture(self.j2_u[c], self.j2_mp[c],
                                                       self.j2_vals[c], self.j2_f[c]))

    def launch(self) -> None:
        for g in self.graphs:
            g.launch()
