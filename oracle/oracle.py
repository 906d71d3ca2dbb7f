"""TEST INFRASTRUCTURE ONLY — ctypes access to the oracle.

  Oracle(kind="port")       liborc.so, our CPU restatement of the reference path
  Oracle(kind="reference")  _ref/libcxpp_ref.so, the reference's calculixpp_core
                            compiled from its own sources (oracle/Makefile `ref`)
  Oracle(kind="dropin")     _ref/libcxpp_dropin.so, the same build with its assembly
                            calls (assemble_linear_static, assemble_material_tangent,
                            internal_force, element_stiffness) routed to the GPU drop-in
                            cxpp::gpu; the model is still the reference's cxpp::Model.
                            Not a checker: what the drop-in tests check.

Both expose the same methods on the same array-level model (orc_model.h). Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PORT_LIB = os.path.join(HERE, "liborc.so")
REF_LIB = os.path.join(HERE, "_ref", "libcxpp_ref.so")
DROPIN_LIB = os.path.join(HERE, "_ref", "libcxpp_dropin.so")
ORC_STATE = 26
MAX_TABLE = 16


class OracleError(RuntimeError):
    pass


class OrcJ2(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("hardening", C.c_int32),
                ("n_table", C.c_int32), ("eqplastic", C.c_double * MAX_TABLE),
                ("yield_", C.c_double * MAX_TABLE), ("kinematic_modulus", C.c_double)]


class OrcModel(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("coords", C.c_void_p), ("n_elem", C.c_int32),
                ("etype", C.c_void_p), ("eptr", C.c_void_p), ("conn", C.c_void_p),
                ("E", C.c_void_p), ("nu", C.c_void_p), ("j2_id", C.c_void_p), ("n_j2", C.c_int32),
                ("j2", C.c_void_p), ("active", C.c_void_p), ("n_spc", C.c_int32),
                ("spc_node", C.c_void_p), ("spc_comp", C.c_void_p), ("spc_value", C.c_void_p),
                ("n_cload", C.c_int32), ("cload_node", C.c_void_p), ("cload_comp", C.c_void_p),
                ("cload_value", C.c_void_p), ("rho", C.c_void_p)]


def build(ref: bool = False) -> None:
    """Compile liborc.so (and, when asked and the reference tree is present, _ref with
    the reference core and the drop-in build; the latter links libcxg.so)."""
    targets = ["all"] + (["ref", "dropin"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE, "-j8"] + targets, check=True)


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def dropin_available() -> bool:
    return os.path.exists(DROPIN_LIB)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def j2_struct(p) -> OrcJ2:
    m = OrcJ2()
    m.E, m.nu, m.hardening = p.E, p.nu, p.hardening
    m.n_table = len(p.yield_)
    for i, (e, y) in enumerate(zip(p.eqplastic, p.yield_)):
        m.eqplastic[i] = e
        m.yield_[i] = y
    m.kinematic_modulus = p.kinematic_modulus
    return m


class _ModelArrays:
    """Keeps the numpy arrays alive behind an OrcModel."""

    def __init__(self, model):
        self.coords = np.ascontiguousarray(model.coords, np.float64)
        self.etype = np.ascontiguousarray(model.element_types(), np.int32)
        self.eptr, self.conn = model.flat_conn()
        self.E = np.ascontiguousarray(model.E, np.float64)
        self.nu = np.ascontiguousarray(model.nu, np.float64)
        self.j2_id = None if model.j2_id is None else np.ascontiguousarray(model.j2_id, np.int32)
        self.j2 = (OrcJ2 * max(len(model.j2), 1))(*[j2_struct(p) for p in model.j2])
        self.active = None if model.active is None else np.ascontiguousarray(model.active, np.uint8)
        self.spc = [np.ascontiguousarray(model.spc_node, np.int32),
                    np.ascontiguousarray(model.spc_comp, np.int32),
                    np.ascontiguousarray(model.spc_value, np.float64)]
        self.cl = [np.ascontiguousarray(model.cload_node, np.int32),
                   np.ascontiguousarray(model.cload_comp, np.int32),
                   np.ascontiguousarray(model.cload_value, np.float64)]
        m = OrcModel()
        m.n_nodes = len(self.coords)
        m.coords = _p(self.coords)
        m.n_elem = len(self.etype)
        m.etype, m.eptr, m.conn = _p(self.etype), _p(self.eptr), _p(self.conn)
        m.E, m.nu = _p(self.E), _p(self.nu)
        m.j2_id = _p(self.j2_id)
        m.n_j2 = len(model.j2)
        m.j2 = C.cast(self.j2, C.c_void_p)
        m.active = _p(self.active)
        m.n_spc = len(self.spc[0])
        m.spc_node, m.spc_comp, m.spc_value = (_p(a) for a in self.spc)
        m.n_cload = len(self.cl[0])
        m.cload_node, m.cload_comp, m.cload_value = (_p(a) for a in self.cl)
        rho = getattr(model, "rho", None)
        self.rho = None if rho is None else np.ascontiguousarray(rho, np.float64)
        m.rho = _p(self.rho)
        self.struct = m


class System:
    """Assembled reduced system: hash-order COO (as flushed), canonical CSR, rhs."""

    def __init__(self, n_free, rows, cols, vals, rhs, dof_eq):
        self.n_free, self.rows, self.cols, self.vals = n_free, rows, cols, vals
        self.rhs, self.dof_eq = rhs, dof_eq

    def csr(self):
        order = np.lexsort((self.cols, self.rows))
        rows, cols, vals = self.rows[order], self.cols[order], self.vals[order]
        row_ptr = np.zeros(self.n_free + 1, np.int64)
        np.add.at(row_ptr, rows.astype(np.int64) + 1, 1)
        return np.cumsum(row_ptr), cols.astype(np.int32), vals

    def dense(self):
        K = np.zeros((self.n_free, self.n_free))
        np.add.at(K, (self.rows, self.cols), self.vals)
        return K


class Oracle:
    def __init__(self, kind: str = "port"):
        self.kind = kind
        path = {"port": PORT_LIB, "reference": REF_LIB, "dropin": DROPIN_LIB}[kind]
        if not os.path.exists(path):
            raise OracleError(f"{path} not built")
        self.L = C.CDLL(path)
        self.pre = "orc_" if kind == "port" else "ref_"
        self._err = C.create_string_buffer(512)

    def _f(self, name):
        return getattr(self.L, self.pre + name)

    def _check(self, st):
        if st != 0:
            raise OracleError(self._err.value.decode())

    # ---- element level -----------------------------------------------------------
    def element_stiffness(self, etype, coords, E, nu):
        x = np.ascontiguousarray(coords, np.float64)
        n = x.size // 3
        Ke = np.zeros((3 * n, 3 * n))
        f = self._f("element_stiffness")
        f.argtypes = [C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_char_p, C.c_int]
        self._check(f(int(etype), _p(x), float(E), float(nu), _p(Ke), self._err, 512))
        return Ke

    def j2_evaluate(self, params, strain, state):
        st = np.ascontiguousarray(state, np.float64).copy()
        eps = np.ascontiguousarray(strain, np.float64)
        sig, D = np.zeros(6), np.zeros((6, 6))
        m = j2_struct(params)
        f = self._f("j2_evaluate")
        f.argtypes = [C.POINTER(OrcJ2), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        self._check(f(C.byref(m), _p(eps), _p(st), _p(sig), _p(D), self._err, 512))
        return sig, D, st

    def element_tangent_force(self, etype, coords, ue, E, nu, params=None, state=None):
        x = np.ascontiguousarray(coords, np.float64)
        u = np.ascontiguousarray(ue, np.float64)
        n = x.size // 3
        Ke, fe = np.zeros((3 * n, 3 * n)), np.zeros(3 * n)
        st = None if state is None else np.ascontiguousarray(state, np.float64).copy()
        m = None if params is None else j2_struct(params)
        f = self._f("element_tangent_force")
        f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.POINTER(OrcJ2),
                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        self._check(f(int(etype), _p(x), _p(u), float(E), float(nu),
                      None if m is None else C.byref(m), _p(st), _p(Ke), _p(fe), self._err, 512))
        return Ke, fe, st

    # ---- model level ---------------------------------------------------------------
    def _system(self, h):
        g = self._f
        g("system_n_free").restype = C.c_int32
        g("system_nnz").restype = C.c_int64
        nf = g("system_n_free")(h)
        nnz = g("system_nnz")(h)
        rows, cols, vals = np.zeros(nnz, np.int32), np.zeros(nnz, np.int32), np.zeros(nnz)
        rhs = np.zeros(nf)
        g("system_coo")(h, _p(rows), _p(cols), _p(vals))
        g("system_rhs")(h, _p(rhs))
        return nf, rows, cols, vals, rhs

    def _dof_eq(self, h, n_nodes):
        d = np.zeros(3 * n_nodes, np.int32)
        self._f("system_dof_eq")(h, _p(d))
        return d

    def open(self, model) -> "ModelHandle":
        """The model built once and kept across calls (a Newton loop's Model)."""
        return ModelHandle(self, model)

    def resident_stats(self, reset: bool = False) -> dict:
        """Residency counters of the drop-in's calling thread (drop-in build only)."""
        out = np.zeros(6, np.int64)
        f = self.L.ref_gpu_resident_stats
        f.argtypes = [C.c_int, C.c_void_p]
        f.restype = None
        f(int(reset), _p(out))
        keys = ("calls", "mesh_uploads", "pattern_builds", "dof_maps", "history_uploads",
                "history_commits")
        return dict(zip(keys, (int(v) for v in out)))

    def solve_sparse(self, n, rows, cols, vals, rhs, direct: bool = False):
        """cxpp::gpu::hip_backend().solve_sparse on COO triplets (drop-in build only):
        (x, backend kind)."""
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        vals = np.ascontiguousarray(vals, np.float64)
        rhs = np.ascontiguousarray(rhs, np.float64)
        x = np.zeros(n)
        kind = C.c_int(-1)
        f = self.L.ref_solve_sparse
        f.argtypes = [C.c_int32, C.c_int64] + [C.c_void_p] * 4 + [C.c_int, C.c_void_p,
                                                                  C.POINTER(C.c_int), C.c_char_p,
                                                                  C.c_int]
        self._check(f(int(n), len(vals), _p(rows), _p(cols), _p(vals), _p(rhs), int(direct),
                      _p(x), C.byref(kind), self._err, 512))
        return x, kind.value

    def _with_model(self, model, fn):
        if isinstance(model, ModelHandle):
            return fn(model.ref, model.arrays)
        arrays = _ModelArrays(model)
        if self.kind == "port":
            return fn(C.byref(arrays.struct), arrays)
        h = C.c_void_p()
        create = self.L.ref_model_create
        create.argtypes = [C.POINTER(OrcModel), C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
        self._check(create(C.byref(arrays.struct), C.byref(h), self._err, 512))
        try:
            return fn(h, arrays)
        finally:
            self.L.ref_model_free(h)

    def assemble_linear_static(self, model) -> System:
        def run(m, arrays):
            h = C.c_void_p()
            f = self._f("assemble_linear_static")
            f.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
            self._check(f(m, C.byref(h), self._err, 512))
            try:
                nf, rows, cols, vals, rhs = self._system(h)
                return System(nf, rows, cols, vals, rhs, self._dof_eq(h, model.n_nodes))
            finally:
                self._f("system_free")(h)
        return self._with_model(model, run)

    # ---- cxpp::gpu::ArrayModel entry points (drop-in build only) ----------------------
    def _array_system(self, model, fn):
        arrays = _ModelArrays(model)
        h = C.c_void_p()
        self._check(fn(C.byref(arrays.struct), C.byref(h)))
        try:
            nf, rows, cols, vals, rhs = self._system(h)
            return System(nf, rows, cols, vals, rhs, self._dof_eq(h, model.n_nodes))
        finally:
            self._f("system_free")(h)

    def array_assemble_linear_static(self, model, via_model: bool = False) -> System:
        """cxpp::gpu::assemble_linear_static on an ArrayModel built straight from the arrays
        (via_model: on cxpp::gpu::to_arrays of the reference Model built from them)."""
        f = self.L.ref_gpu_array_linear_static
        f.argtypes = [C.POINTER(OrcModel), C.c_int, C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
        return self._array_system(model, lambda m, h: f(m, 0 if via_model else 1, h, self._err, 512))

    def array_assemble_mass(self, model, lumped: bool = False) -> System:
        f = self.L.ref_gpu_array_mass
        f.argtypes = [C.POINTER(OrcModel), C.c_int, C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
        return self._array_system(model, lambda m, h: f(m, int(lumped), h, self._err, 512))

    def array_internal_force(self, model, u):
        arrays = _ModelArrays(model)
        uu = np.ascontiguousarray(u, np.float64).ravel()
        out = np.zeros(3 * model.n_nodes)
        f = self.L.ref_gpu_array_internal_force
        f.argtypes = [C.POINTER(OrcModel), C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        self._check(f(C.byref(arrays.struct), _p(uu), _p(out), self._err, 512))
        return out

    def assemble_material_tangent(self, model, u, state):
        """state: (n_points, 26) AoS history, updated in place copy; returns
        (System, f_int, new_state)."""
        st = np.ascontiguousarray(state, np.float64).copy()
        uu = np.ascontiguousarray(u, np.float64).reshape(-1)
        fi = np.zeros(3 * model.n_nodes)

        def run(m, arrays):
            h = C.c_void_p()
            f = self._f("assemble_material_tangent")
            f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p),
                          C.c_char_p, C.c_int]
            self._check(f(m, _p(uu), _p(st), _p(fi), C.byref(h), self._err, 512))
            try:
                nf, rows, cols, vals, rhs = self._system(h)
                return System(nf, rows, cols, vals, rhs, self._dof_eq(h, model.n_nodes))
            finally:
                self._f("system_free")(h)
        sys = self._with_model(model, run)
        return sys, fi, st

    def internal_force(self, model, u):
        uu = np.ascontiguousarray(u, np.float64).reshape(-1)
        fi = np.zeros(3 * model.n_nodes)

        def run(m, arrays):
            f = self._f("internal_force")
            f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
            self._check(f(m, _p(uu), _p(fi), self._err, 512))
            return fi
        return self._with_model(model, run)

    def time_linear_static(self, model, u=None):
        """Seconds spent in assemble_linear_static (+ internal_force when u is given)
        alone: model construction and result copies are outside the clock."""
        import time
        uu = None if u is None else np.ascontiguousarray(u, np.float64).reshape(-1)
        fi = np.zeros(3 * model.n_nodes)

        def run(m, arrays):
            h = C.c_void_p()
            f = self._f("assemble_linear_static")
            f.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
            g = self._f("internal_force")
            g.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
            t0 = time.perf_counter()
            st = f(m, C.byref(h), self._err, 512)
            if st == 0 and uu is not None:
                st = g(m, _p(uu), _p(fi), self._err, 512)
            t1 = time.perf_counter()
            if h.value:
                self._f("system_free")(h)
            self._check(st)
            return t1 - t0
        return self._with_model(model, run)

    def time_material_tangent(self, model, u, state):
        """Seconds spent in assemble_material_tangent alone (state: (n_points, 26) AoS,
        copied first; model construction and result copies outside the clock)."""
        import time
        st = np.ascontiguousarray(state, np.float64).copy()
        uu = np.ascontiguousarray(u, np.float64).reshape(-1)
        fi = np.zeros(3 * model.n_nodes)

        def run(m, arrays):
            h = C.c_void_p()
            f = self._f("assemble_material_tangent")
            f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p),
                          C.c_char_p, C.c_int]
            t0 = time.perf_counter()
            status = f(m, _p(uu), _p(st), _p(fi), C.byref(h), self._err, 512)
            t1 = time.perf_counter()
            if h.value:
                self._f("system_free")(h)
            self._check(status)
            return t1 - t0
        return self._with_model(model, run)

    def _assemble_scalar(self, model, kind, gp_stress=None) -> System:
        g = None if gp_stress is None else np.ascontiguousarray(gp_stress, np.float64)

        def run(m, arrays):
            h = C.c_void_p()
            f = self._f("assemble_scalar")
            f.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_void_p), C.c_char_p,
                          C.c_int]
            self._check(f(m, int(kind), None if g is None else _p(g), C.byref(h), self._err, 512))
            try:
                nf, rows, cols, vals, rhs = self._system(h)
                return System(nf, rows, cols, vals, rhs, self._dof_eq(h, model.n_nodes))
            finally:
                self._f("system_free")(h)
        return self._with_model(model, run)

    def assemble_mass(self, model, lumped=False) -> System:
        """assemble_mass (consistent, or row-sum lumped) with model.rho."""
        return self._assemble_scalar(model, 1 if lumped else 0)

    def assemble_geometric_stiffness(self, model, gp_stress) -> System:
        """assemble_geometric_stiffness; gp_stress (n_points, 6) in model element order
        (rows of removed elements are ignored)."""
        return self._assemble_scalar(model, 2, gp_stress)

    def recover_fields(self, model, u, state=None, gauss_stress=False):
        """recover_fields (state None) or its material-point form (state: (n_points, 26)
        AoS history, committed slots read): dict of strain, stress (n, 6), reaction,
        f_int (n, 3) and, with gauss_stress, the per-point stress (n_points, 6) in
        model element order (zeros for removed elements). For the reference, f_int is
        reaction + f_ext and gauss_stress exists for state None only."""
        uu = np.ascontiguousarray(u, np.float64).reshape(-1)
        st = None if state is None else np.ascontiguousarray(state, np.float64).copy()
        n = model.n_nodes
        out = {"strain": np.zeros((n, 6)), "stress": np.zeros((n, 6)),
               "reaction": np.zeros((n, 3)), "f_int": np.zeros((n, 3))}
        gs = np.zeros((max(model.n_points, 1), 6)) if gauss_stress else None

        def run(m, arrays):
            f = self._f("recover_fields")
            f.argtypes = [C.c_void_p] + [C.c_void_p] * 7 + [C.c_char_p, C.c_int]
            self._check(f(m, _p(uu), None if st is None else _p(st), _p(out["strain"]),
                          _p(out["stress"]), _p(out["reaction"]), _p(out["f_int"]),
                          None if gs is None else _p(gs), self._err, 512))
        self._with_model(model, run)
        if gs is not None:
            out["gauss_stress"] = gs[: model.n_points]
        return out

    def vjp(self, model, u, w):
        """(dx (n,3), dE (n_elem,), du (3n,), r (3n,)) of r = K(x,E)u, port only."""
        if self.kind != "port":
            raise OracleError("vjp exists in the restatement only")
        arrays = _ModelArrays(model)
        uu = np.ascontiguousarray(u, np.float64).reshape(-1)
        ww = np.ascontiguousarray(w, np.float64).reshape(-1)
        dx, du, r = np.zeros(3 * model.n_nodes), np.zeros(3 * model.n_nodes), np.zeros(3 * model.n_nodes)
        dE = np.zeros(model.n_elem)
        f = self.L.orc_vjp
        f.argtypes = [C.c_void_p] + [C.c_void_p] * 6 + [C.c_char_p, C.c_int]
        self._check(f(C.byref(arrays.struct), _p(uu), _p(ww), _p(dx), _p(dE), _p(du), _p(r),
                      self._err, 512))
        return dx.reshape(-1, 3), dE, du, r


class ModelHandle:
    """One model built once by an Oracle and passed to its methods in place of the array
    model: for the reference and drop-in libraries a cxpp::Model that lives until close(),
    as the Model a Newton loop holds across its calls. Attributes of the array model
    (n_nodes, n_points, ...) read through."""

    def __init__(self, oracle: Oracle, model):
        self.model = model
        self.arrays = _ModelArrays(model)
        self._oracle = oracle
        if oracle.kind == "port":
            self.ref = C.byref(self.arrays.struct)
            return
        h = C.c_void_p()
        create = oracle.L.ref_model_create
        create.argtypes = [C.POINTER(OrcModel), C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
        oracle._check(create(C.byref(self.arrays.struct), C.byref(h), oracle._err, 512))
        self.ref = h

    def __getattr__(self, name):
        return getattr(self.model, name)

    def close(self):
        if self._oracle.kind != "port" and self.ref is not None and self.ref.value:
            self._oracle.L.ref_model_free(self.ref)
        self.ref = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def aos_state(soa: np.ndarray) -> np.ndarray:
    """(13, P) committed SoA -> (P, 26) oracle AoS with trial = committed."""
    P = soa.shape[1]
    st = np.zeros((P, ORC_STATE))
    st[:, :13] = soa.T
    st[:, 13:] = soa.T
    return st
