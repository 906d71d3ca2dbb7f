"""Field recovery on the device: nodal stress and strain, internal force and reactions.

  recover_fields(model, u)           fem::recover_fields (src/fem/stress.cpp:189-273):
                                     every element elastic (D eps), f_int = sum_e Ke ue
  recover_fields(model, u, mp)       its material-point form (stress.cpp:275-353): J2
                                     elements take the radial-return stress from the
                                     committed history, f_int from element_tangent_force
  recover_gauss_stress(model, u)     fem::recover_gauss_stress (stress.cpp:147-187):
                                     per element, the stress at each Gauss point

All run as cxg_recover_fields: one wave per element for the Gauss-point fields and the
element means, one thread per node for the averages over the node's elements (ascending
element order) and RF = f_int - f_ext. The hex and wedge families extrapolate by the
Gauss-point mean at every node (fem/extrapolation.hpp:37-45).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import torch

from . import _lib, fem
from .model import GAUSS, Model


@dataclass
class StaticFields:
    """core/results.hpp:26-35 (without the contact results)."""
    displacement: np.ndarray  # (n_nodes, 3)
    stress: np.ndarray        # (n_nodes, 6) averaged nodal stress, Voigt xx,yy,zz,xy,xz,yz
    strain: np.ndarray        # (n_nodes, 6) averaged nodal strain, engineering shear
    reaction: np.ndarray      # (n_nodes, 3) f_int - f_ext


class FieldRecovery:
    """The mesh (active elements only) and the material table kept resident, for
    repeated recoveries on one model."""

    def __init__(self, model: Model, ctx: Optional[fem.Context] = None):
        self.model = model
        self.active = None
        if model.active is not None and not np.all(model.active):
            self.active = np.asarray(model.active, bool)
            model = fem._only_active(model)  # removed elements add no stress (stress.cpp:204)
        self.mesh = fem.DeviceMesh(model, ctx)
        self.ctx = self.mesh.ctx
        dev = self.ctx.device
        self.n_points = self.mesh.n_points
        self.f_ext = torch.from_numpy(self.model.external_load()).to(dev)
        self.mats = (_lib.J2Material * max(len(self.model.j2), 1))()
        for i, p in enumerate(self.model.j2):
            self.mats[i] = p.to_c()
        self.n_mats = len(self.model.j2)
        jid = self.model.j2_id
        if jid is not None and self.active is not None:
            jid = jid[self.active]
        self.mat_id = torch.from_numpy(
            jid if jid is not None else np.full(self.mesh.n_elem, -1, np.int32)).to(dev)

    def run(self, u: torch.Tensor, mp: Optional[fem.MaterialPoints] = None,
            gauss_stress: bool = False, nodal: bool = True) -> dict:
        """Device tensors: strain, stress (n_nodes, 6), f_int, reaction (n_nodes, 3) and,
        with gauss_stress, the per-point stress (n_points, 6) of the active elements."""
        dev = self.ctx.device
        n = self.model.n_nodes
        u = u.to(dev, torch.float64).contiguous()
        out = {}
        if nodal:
            out["strain"] = torch.empty((n, 6), dtype=torch.float64, device=dev)
            out["stress"] = torch.empty((n, 6), dtype=torch.float64, device=dev)
            out["f_int"] = torch.empty((n, 3), dtype=torch.float64, device=dev)
            out["reaction"] = torch.empty((n, 3), dtype=torch.float64, device=dev)
        if gauss_stress:
            out["gauss_stress"] = torch.empty((max(self.n_points, 1), 6), dtype=torch.float64,
                                              device=dev)
        with_mp = mp is not None and self.n_mats > 0
        if mp is not None and self.active is not None:
            raise NotImplementedError("*MODEL CHANGE on the material-point path")
        m = self.mesh
        p = fem._ptr
        self.ctx.check(self.ctx.lib.cxg_recover_fields(
            self.ctx.h, m.h, p(m.coords), p(m.E), p(m.nu), p(self.mat_id),
            self.mats, self.n_mats if with_mp else 0, p(u), p(mp.committed) if with_mp else None,
            p(self.f_ext), p(out.get("strain")), p(out.get("stress")), p(out.get("reaction")),
            p(out.get("f_int")), p(out.get("gauss_stress"))))
        return out


def recover_fields(model: Model, displacement, mp: Optional[fem.MaterialPoints] = None,
                   device=None) -> StaticFields:
    """StaticFields of a solved displacement (n_nodes, 3): nodal stress and strain and
    the reactions; with mp (committed J2 history) the material-point form."""
    r = FieldRecovery(model, fem.context(device))
    u = displacement if torch.is_tensor(displacement) else torch.from_numpy(
        np.ascontiguousarray(np.asarray(displacement, np.float64).reshape(-1, 3)))
    o = r.run(u, mp)
    return StaticFields(u.reshape(-1, 3).cpu().numpy(), o["stress"].cpu().numpy(),
                        o["strain"].cpu().numpy(), o["reaction"].cpu().numpy())


def recover_gauss_stress(model: Model, u, device=None) -> List[np.ndarray]:
    """Per model element, its Gauss-point stress (n_gp, 6); removed elements get an
    empty (0, 6) array."""
    r = FieldRecovery(model, fem.context(device))
    uu = u if torch.is_tensor(u) else torch.from_numpy(
        np.ascontiguousarray(np.asarray(u, np.float64).reshape(-1, 3)))
    gs = r.run(uu, gauss_stress=True, nodal=False)["gauss_stress"].cpu().numpy()
    out: List[np.ndarray] = []
    off = 0
    e = 0
    for t, c in model.blocks:
        ng = GAUSS[t]
        for _ in range(c.shape[0]):
            if r.active is not None and not r.active[e]:
                out.append(np.zeros((0, 6)))
            else:
                out.append(gs[off: off + ng].copy())
                off += ng
            e += 1
    return out
