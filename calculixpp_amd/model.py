"""Array-native model description for the GPU assembly path.

The reference builds a `cxpp::Model` (include/calculixpp/core/model.hpp:294) from an
.inp deck: node ids, elements with node ids, *SOLID SECTION -> *MATERIAL via elsets,
*BOUNDARY (Spc, model.hpp:22) and *CLOAD (Cload, model.hpp:44). At the sizes of this
path (262k elements with one E each) that is one material and elset per element, so
the model here is the same information as arrays:

  coords   (n_nodes, 3) float64       node coordinates, node INDEX order (ids are
                                      index+1 in the reference's insertion order)
  blocks   [(type, conn (n_e, nn) int32)] element runs of one family; model element
                                      numbers follow block order
  E, nu    (n_elem,) float64          per-element isotropic elasticity
  spc_*    node index, component 1..3, value     (*BOUNDARY)
  cload_*  node index, component 1..3, value     (*CLOAD)
  j2_id    (n_elem,) int32, -1 = elastic; j2 = list of J2Params  (*PLASTIC)
  active   (n_elem,) bool or None                (*MODEL CHANGE, REMOVE)
  rho      (n_elem,) float64, default 0          (*DENSITY, Model::element_density)
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib

NODES = {_lib.C3D8: 8, _lib.C3D20: 20, _lib.C3D20R: 20, _lib.C3D6: 6}
GAUSS = {_lib.C3D8: 8, _lib.C3D20: 27, _lib.C3D20R: 8, _lib.C3D6: 2}


@dataclass
class J2Params:
    """cxpp::Plastic + ElasticIso (core/material.hpp:11-44)."""
    E: float
    nu: float
    eqplastic: list
    yield_: list
    hardening: int = _lib.HARD_ISOTROPIC
    kinematic_modulus: float = 0.0

    def to_c(self) -> _lib.J2Material:
        if not 1 <= len(self.yield_) <= _lib.MAX_TABLE or len(self.yield_) != len(self.eqplastic):
            raise ValueError("J2PlasticMaterial needs a non-empty *PLASTIC table")
        m = _lib.J2Material()
        m.E, m.nu, m.hardening = self.E, self.nu, self.hardening
        m.n_table = len(self.yield_)
        for i, (e, y) in enumerate(zip(self.eqplastic, self.yield_)):
            m.eqplastic[i] = e
            m.yield_[i] = y
        m.kinematic_modulus = self.kinematic_modulus
        return m


@dataclass
class Model:
    coords: np.ndarray
    blocks: list
    E: np.ndarray
    nu: np.ndarray
    spc_node: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    spc_comp: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    spc_value: np.ndarray = field(default_factory=lambda: np.zeros(0))
    cload_node: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    cload_comp: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    cload_value: np.ndarray = field(default_factory=lambda: np.zeros(0))
    j2_id: Optional[np.ndarray] = None
    j2: list = field(default_factory=list)
    active: Optional[np.ndarray] = None
    rho: Optional[np.ndarray] = None

    def __post_init__(self):
        self.coords = np.ascontiguousarray(self.coords, np.float64)
        self.blocks = [(int(t), np.ascontiguousarray(c, np.int32)) for t, c in self.blocks]
        for t, c in self.blocks:
            if t not in NODES:
                raise ValueError(f"element type {t} outside the GPU path")
            if c.ndim != 2 or c.shape[1] != NODES[t]:
                raise ValueError("connectivity shape does not match the element type")
            srt = np.sort(c, axis=1)
            if c.shape[0] and bool((srt[:, 1:] == srt[:, :-1]).any()):
                # one incidence per node and element is what the row gather and the
                # node sums assume; a collapsed element repeats a node
                raise ValueError("an element lists a node twice (collapsed element): "
                                 "outside the GPU path")
        n = self.n_elem
        self.E = np.array(np.broadcast_to(np.asarray(self.E, np.float64), (n,)))
        self.nu = np.array(np.broadcast_to(np.asarray(self.nu, np.float64), (n,)))
        self.rho = np.zeros(n) if self.rho is None else \
            np.array(np.broadcast_to(np.asarray(self.rho, np.float64), (n,)))
        for name in ("spc_node", "spc_comp", "cload_node", "cload_comp"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name), np.int32))
        for name in ("spc_value", "cload_value"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name), np.float64))
        if self.j2_id is not None:
            self.j2_id = np.ascontiguousarray(self.j2_id, np.int32)

    @property
    def n_nodes(self) -> int:
        return int(self.coords.shape[0])

    @property
    def n_elem(self) -> int:
        return int(sum(c.shape[0] for _, c in self.blocks))

    @property
    def n_points(self) -> int:
        return int(sum(c.shape[0] * GAUSS[t] for t, c in self.blocks))

    def element_types(self) -> np.ndarray:
        return np.concatenate([np.full(c.shape[0], t, np.int32) for t, c in self.blocks])

    def flat_conn(self):
        """(eptr int64[n_elem+1], conn int32[...]) over all blocks, model order."""
        sizes = np.concatenate([np.full(c.shape[0], c.shape[1], np.int64) for _, c in self.blocks])
        eptr = np.zeros(len(sizes) + 1, np.int64)
        np.cumsum(sizes, out=eptr[1:])
        conn = np.concatenate([c.ravel() for _, c in self.blocks]).astype(np.int32)
        return eptr, conn

    def dof_map(self):
        """SPC-only free numbering (fem/assembly.cpp:19-42, constraint_transform.cpp:
        78-82): (dof_eq int32[3n], n_free, prescribed float64[3n])."""
        nd = 3 * self.n_nodes
        fixed = np.zeros(nd, bool)
        pres = np.zeros(nd)
        if len(self.spc_node):
            if self.spc_node.min() < 0 or self.spc_node.max() >= self.n_nodes:
                raise RuntimeError("*BOUNDARY references unknown node")
            if self.spc_comp.min() < 1 or self.spc_comp.max() > 3:
                raise RuntimeError("*BOUNDARY dof out of range")
            g = 3 * self.spc_node.astype(np.int64) + self.spc_comp - 1
            fixed[g] = True
            # the last card of a DOF wins, as in the reference's sequential loop: the
            # first occurrence of each DOF in the reversed card list
            gr = g[::-1]
            uniq, first = np.unique(gr, return_index=True)
            pres[uniq] = self.spc_value[::-1][first]
        dof_eq = np.full(nd, -1, np.int32)
        free = ~fixed
        dof_eq[free] = np.arange(int(free.sum()), dtype=np.int32)
        return dof_eq, int(free.sum()), pres

    def external_load(self) -> np.ndarray:
        """Full-DOF *CLOAD vector (fem/loads.cpp:579-585, lambda = 1)."""
        f = np.zeros(3 * self.n_nodes)
        ok = (self.cload_node >= 0) & (self.cload_node < self.n_nodes) & \
             (self.cload_comp >= 1) & (self.cload_comp <= 3)
        np.add.at(f, 3 * self.cload_node[ok].astype(np.int64) + self.cload_comp[ok] - 1,
                  self.cload_value[ok])
        return f
