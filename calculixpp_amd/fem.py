"""Python mirror of the reference's fem:: assembly interface, on the HIP path.

Same names, argument meaning and error behaviour as the reference's C++ functions
(include/calculixpp/fem/assembly.hpp, element.hpp, stress.hpp), with device
arrays (torch tensors on the current HIP device) in place of std::vector:

  assemble_linear_static(model)                 fem/assembly.cpp:225
  make_material_points(model) / .commit()       fem/assembly.cpp:357, assembly.hpp:84
  assemble_material_tangent(model, u, mp)       fem/assembly.cpp:385  -> (LinearSystem, f_int)
  internal_force(model, u)                      fem/stress.cpp:125
  element_stiffness(type, coords, E, nu)        fem/element.cpp:614
  assemble_mass(model, lumped)                  fem/assembly.cpp:259
  assemble_geometric_stiffness(model, gp_stress) fem/assembly.cpp:299

Field recovery (recover_fields, recover_gauss_stress) lives in calculixpp_amd.recover.

The returned LinearSystem holds the free/free stiffness as canonical CSR (rows
ascending, columns ascending), which is what SciPP's CsrMatrix::from_coo builds from
the reference's COO (src/compute/cpu_backend.cpp:46); .coo() gives the reference's
rows/cols/vals triplets. Errors raise CxgError (a RuntimeError) with the reference's
message text.

For repeated assemblies on one mesh (Newton loops, benchmarks) use Assembler, which
uploads the mesh and builds the pattern once.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib
from .model import GAUSS, NODES, Model


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


def _dev(device=None) -> torch.device:
    if device is None:
        if not torch.cuda.is_available():
            raise _lib.CxgLibraryError("the HIP path needs a GPU (torch.cuda.is_available() is False)")
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


class Context:
    """A cxg_context bound to a HIP device and the current torch stream."""

    def __init__(self, device=None):
        self.device = _dev(device)
        self.lib = _lib.lib()
        h = C.c_void_p()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(None, self.lib.cxg_create(self.device.index or 0, C.c_void_p(stream), C.byref(h)))
        self.h = h
        self.options = {"asm_variant": 7, "patch_size": 8}

    def sync_stream(self):
        stream = torch.cuda.current_stream(self.device).cuda_stream
        self.lib.cxg_set_stream(self.h, C.c_void_p(stream))

    def check(self, status: int):
        _lib.check(self.h, status)

    def check_kernels(self) -> None:
        """cxg_check: synchronize and raise kernel status flags (after graph replays)."""
        self.check(self.lib.cxg_check(self.h))

    def set_option(self, key: str, value: int) -> None:
        """cxg_set_option: "patch_size" (meshes created afterwards), "asm_variant"."""
        self.check(self.lib.cxg_set_option(self.h, key.encode(), int(value)))
        self.options[key] = int(value)

    @property
    def tracks_live(self) -> bool:
        """Whether assemblies can report entry liveness (the two-phase kernels)."""
        return self.options.get("asm_variant", 7) >= 4

    def profile(self, enable: bool) -> None:
        """Bracket every library kernel launch with HIP events on its stream."""
        self.check(self.lib.cxg_profile_enable(self.h, int(enable)))

    def profile_reset(self) -> None:
        self.check(self.lib.cxg_profile_reset(self.h))

    def profile_stats(self) -> dict:
        """{kernel name: (launches, total ms, elements)} accumulated while enabled."""
        out = {}
        for i in range(self.lib.cxg_profile_count(self.h)):
            name, n, ms, el = C.c_char_p(), C.c_int64(), C.c_double(), C.c_int64()
            self.check(self.lib.cxg_profile_get(self.h, i, C.byref(name), C.byref(n),
                                                C.byref(ms), C.byref(el)))
            out[name.value.decode()] = (n.value, ms.value, el.value)
        return out

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self.lib.cxg_destroy(h)
            except Exception:  # pragma: no cover - interpreter shutdown
                pass
            self.h = None


_CTX: dict = {}


def context(device=None) -> Context:
    dev = _dev(device)
    ctx = _CTX.get(dev.index)
    if ctx is None:
        ctx = _CTX[dev.index] = Context(dev)
    ctx.sync_stream()
    return ctx


class DeviceMesh:
    """Mesh arrays resident in HBM + the library's mesh object (colouring)."""

    def __init__(self, model: Model, ctx: Optional[Context] = None):
        self.ctx = ctx or context()
        dev = self.ctx.device
        self.model = model
        self.n_nodes = model.n_nodes
        self.coords = torch.from_numpy(model.coords).to(dev)
        self.E = torch.from_numpy(model.E).to(dev)
        self.nu = torch.from_numpy(model.nu).to(dev)
        self.conn = [torch.from_numpy(c).to(dev) for _, c in model.blocks]
        blocks = (_lib.Block * max(len(model.blocks), 1))()
        eoff = goff = 0
        for i, ((t, c), dc) in enumerate(zip(model.blocks, self.conn)):
            blocks[i].type = t
            blocks[i].n_elem = c.shape[0]
            blocks[i].elem_offset = eoff
            blocks[i].gp_offset = goff
            blocks[i].conn = dc.data_ptr()
            eoff += c.shape[0]
            goff += c.shape[0] * GAUSS[t]
        self.n_elem, self.n_points = eoff, goff
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.cxg_mesh_create(self.ctx.h, self.n_nodes, len(model.blocks),
                                                    blocks, C.byref(h)))
        self.h = h
        nc = C.c_int32()
        self.ctx.lib.cxg_mesh_info(h, None, None, C.byref(nc))
        self.n_colors = nc.value

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.ctx.lib.cxg_mesh_destroy(h)
            self.h = None


class Pattern:
    """CSR pattern (row_ptr int64, col_idx int32) + element -> slot map."""

    def __init__(self, mesh: DeviceMesh, dof_eq: np.ndarray, n_free: int):
        self.mesh, self.ctx = mesh, mesh.ctx
        self.dof_eq = torch.from_numpy(np.ascontiguousarray(dof_eq, np.int32)).to(self.ctx.device)
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.cxg_pattern_build(self.ctx.h, mesh.h, _ptr(self.dof_eq),
                                                      int(n_free), C.byref(h)))
        self.h = h
        nf, nnz = C.c_int32(), C.c_int64()
        self.ctx.lib.cxg_pattern_info(h, C.byref(nf), C.byref(nnz))
        self.n_free, self.nnz = nf.value, nnz.value
        self._row_ptr = self._col_idx = None

    @property
    def row_ptr(self) -> torch.Tensor:
        if self._row_ptr is None:
            self._row_ptr = torch.empty(self.n_free + 1, dtype=torch.int64, device=self.ctx.device)
            self.ctx.check(self.ctx.lib.cxg_pattern_copy(self.ctx.h, self.h, _ptr(self._row_ptr), None))
        return self._row_ptr

    @property
    def col_idx(self) -> torch.Tensor:
        if self._col_idx is None:
            self._col_idx = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=self.ctx.device)
            self.ctx.check(self.ctx.lib.cxg_pattern_copy(self.ctx.h, self.h, None, _ptr(self._col_idx)))
        return self._col_idx[: self.nnz]

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.ctx.lib.cxg_pattern_destroy(h)
            self.h = None


@dataclass
class LinearSystem:
    """Reduced K_ff (canonical CSR on the device) + rhs (fem/assembly.hpp:18-38)."""
    n_free: int
    row_ptr: torch.Tensor
    col_idx: torch.Tensor
    vals: torch.Tensor
    rhs: torch.Tensor
    dof_eq: np.ndarray
    prescribed: np.ndarray

    def coo(self):
        """(rows, cols, vals) numpy triplets, as LinearSystem.rows/cols/vals."""
        rp = self.row_ptr.cpu().numpy()
        rows = np.repeat(np.arange(self.n_free, dtype=np.int32), np.diff(rp))
        return rows, self.col_idx.cpu().numpy(), self.vals.cpu().numpy()

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.vals.cpu().numpy(), self.col_idx.cpu().numpy(),
                              self.row_ptr.cpu().numpy()), shape=(self.n_free, self.n_free))


class Assembler:
    """Mesh + pattern uploaded once; assembles K (and optionally f_int) many times."""

    def __init__(self, model: Model, ctx: Optional[Context] = None):
        if model.active is not None and not np.all(model.active):
            model = _only_active(model)
        self.model = model
        self.mesh = DeviceMesh(model, ctx)
        self.ctx = self.mesh.ctx
        self.dof_eq, self.n_free, self.prescribed = model.dof_map()
        self.pattern = Pattern(self.mesh, self.dof_eq, self.n_free)

    def new_values(self) -> torch.Tensor:
        return torch.empty(max(self.pattern.nnz, 1), dtype=torch.float64, device=self.ctx.device)

    def new_live(self) -> torch.Tensor:
        """Per-entry flags: 1 where some element contributed a value other than 0.0."""
        return torch.empty(max(self.pattern.nnz, 1), dtype=torch.uint8, device=self.ctx.device)

    def assemble(self, vals: torch.Tensor, u: Optional[torch.Tensor] = None,
                 f_int: Optional[torch.Tensor] = None, coords: Optional[torch.Tensor] = None,
                 E: Optional[torch.Tensor] = None, live: Optional[torch.Tensor] = None) -> None:
        m = self.mesh
        self.ctx.check(self.ctx.lib.cxg_assemble_linear(
            self.ctx.h, self.pattern.h, _ptr(coords if coords is not None else m.coords),
            _ptr(E if E is not None else m.E), _ptr(m.nu), _ptr(vals), _ptr(u), _ptr(f_int),
            _ptr(live)))

    def assemble_mass(self, vals: torch.Tensor, lumped: bool = False,
                      rho: Optional[torch.Tensor] = None, live: Optional[torch.Tensor] = None) -> None:
        """Consistent (or row-sum lumped) mass into vals on this pattern (cxg_assemble_mass);
        rho per element defaults to the model's densities."""
        m = self.mesh
        if rho is None:
            rho = torch.from_numpy(self.model.rho).to(self.ctx.device)
        self.ctx.check(self.ctx.lib.cxg_assemble_mass(self.ctx.h, self.pattern.h, _ptr(m.coords),
                                                      _ptr(rho), int(bool(lumped)), _ptr(vals),
                                                      _ptr(live)))

    def assemble_geometric(self, vals: torch.Tensor, gp_stress: torch.Tensor,
                           live: Optional[torch.Tensor] = None) -> None:
        """Geometric stiffness of the Gauss-point stress (n_points, 6) into vals
        (cxg_assemble_geometric)."""
        m = self.mesh
        g = gp_stress.to(self.ctx.device, torch.float64).contiguous()
        if g.shape != (m.n_points, 6):
            raise ValueError(f"gp_stress must be ({m.n_points}, 6)")
        self.ctx.check(self.ctx.lib.cxg_assemble_geometric(self.ctx.h, self.pattern.h,
                                                           _ptr(m.coords), _ptr(g), _ptr(vals),
                                                           _ptr(live)))

    def same_direction(self):
        """(row_ptr, col_idx) of the same-direction pattern (row (a, i): the structural
        columns (b, i)), which holds the key set of k_ab I3 block matrices; built once."""
        if getattr(self, "_sd", None) is None:
            n = C.c_int64()
            self.ctx.check(self.ctx.lib.cxg_pattern_same_direction(
                self.ctx.h, self.pattern.h, None, None, C.byref(n)))
            dev = self.ctx.device
            rp = torch.empty(self.n_free + 1, dtype=torch.int64, device=dev)
            ci = torch.empty(max(n.value, 1), dtype=torch.int32, device=dev)
            self.ctx.check(self.ctx.lib.cxg_pattern_same_direction(
                self.ctx.h, self.pattern.h, _ptr(rp), _ptr(ci), C.byref(n)))
            self._sd = (rp, ci[: n.value])
        return self._sd

    def sd_fits(self) -> bool:
        """Whether the same-direction gather fits this pattern (every row within the
        gather's LDS columns; else cxg_pattern_same_direction reports UNSUPPORTED)."""
        try:
            self.same_direction()
        except _lib.CxgError as e:
            if e.status != _lib.ERR_UNSUPPORTED:
                raise
            return False
        return True

    def assemble_mass_sd(self, vals: torch.Tensor, live: torch.Tensor, lumped: bool = False,
                         rho: Optional[torch.Tensor] = None) -> None:
        """assemble_mass with vals and live on the same-direction pattern
        (cxg_assemble_mass_sd)."""
        m = self.mesh
        if rho is None:
            rho = torch.from_numpy(self.model.rho).to(self.ctx.device)
        self.same_direction()
        self.ctx.check(self.ctx.lib.cxg_assemble_mass_sd(
            self.ctx.h, self.pattern.h, _ptr(m.coords), _ptr(rho), int(bool(lumped)), _ptr(vals),
            _ptr(live)))

    def assemble_geometric_sd(self, vals: torch.Tensor, live: torch.Tensor,
                              gp_stress: torch.Tensor) -> None:
        """assemble_geometric on the same-direction pattern (cxg_assemble_geometric_sd)."""
        m = self.mesh
        g = gp_stress.to(self.ctx.device, torch.float64).contiguous()
        if g.shape != (m.n_points, 6):
            raise ValueError(f"gp_stress must be ({m.n_points}, 6)")
        self.same_direction()
        self.ctx.check(self.ctx.lib.cxg_assemble_geometric_sd(
            self.ctx.h, self.pattern.h, _ptr(m.coords), _ptr(g), _ptr(vals), _ptr(live)))

    def sd_system(self, vals: torch.Tensor, live: torch.Tensor):
        """(row_ptr, col_idx, vals) of the key set from same-direction values: the pattern
        itself when every entry is live, else its live entries (cxg_compact_csr)."""
        rp, ci = self.same_direction()
        nnz = ci.numel()
        va = vals[:nnz]
        if nnz == 0 or bool(live[:nnz].all()):
            return rp, ci, va
        dev = self.ctx.device
        rp_o = torch.empty_like(rp)
        ci_o = torch.empty(nnz, dtype=torch.int32, device=dev)
        va_o = torch.empty(nnz, dtype=torch.float64, device=dev)
        kept = C.c_int64()
        self.ctx.check(self.ctx.lib.cxg_compact_csr(self.ctx.h, self.n_free, _ptr(rp), _ptr(ci),
                                                    _ptr(live), _ptr(va), _ptr(rp_o), _ptr(ci_o),
                                                    _ptr(va_o), C.byref(kept)))
        return rp_o, ci_o[: kept.value], va_o[: kept.value]

    def capture(self, vals: torch.Tensor, u: Optional[torch.Tensor] = None,
                f_int: Optional[torch.Tensor] = None) -> Graph:
        """One assembly (as assemble with these buffers) captured as a HIP graph."""
        m = self.mesh
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.cxg_graph_capture_linear(
            self.ctx.h, self.pattern.h, _ptr(m.coords), _ptr(m.E), _ptr(m.nu), _ptr(vals),
            _ptr(u), _ptr(f_int), C.byref(h)))
        return Graph(self.ctx, h, (self, vals, u, f_int))

    def rhs(self) -> torch.Tensor:
        """Reduced load: CLOADs through the transform minus the prescribed columns."""
        dev = self.ctx.device
        rhs = torch.zeros(max(self.n_free, 1), dtype=torch.float64, device=dev)
        f_ext = torch.from_numpy(self.model.external_load()).to(dev)
        f_c = None
        if np.any(self.prescribed != 0.0):
            u_c = torch.from_numpy(self.prescribed.reshape(-1, 3).copy()).to(dev)
            f_c = torch.empty(3 * self.model.n_nodes, dtype=torch.float64, device=dev)
            self.ctx.check(self.ctx.lib.cxg_internal_force(
                self.ctx.h, self.mesh.h, _ptr(self.mesh.coords), _ptr(self.mesh.E),
                _ptr(self.mesh.nu), _ptr(u_c), _ptr(f_c)))
        self.ctx.check(self.ctx.lib.cxg_reduce_free(self.ctx.h, self.model.n_nodes,
                                                    _ptr(self.pattern.dof_eq), _ptr(f_ext),
                                                    _ptr(f_c), _ptr(rhs)))
        return rhs[: self.n_free]

    def compact(self, vals: torch.Tensor, live: torch.Tensor):
        """(row_ptr, col_idx, vals) holding only the live entries: the key set of the
        reference's COO map (scatter_block skips 0.0 contributions, assembly.cpp:90)."""
        dev = self.ctx.device
        nnz = max(self.pattern.nnz, 1)
        rp = torch.empty(self.n_free + 1, dtype=torch.int64, device=dev)
        ci = torch.empty(nnz, dtype=torch.int32, device=dev)
        va = torch.empty(nnz, dtype=torch.float64, device=dev)
        kept = C.c_int64()
        self.ctx.check(self.ctx.lib.cxg_pattern_compact(self.ctx.h, self.pattern.h, _ptr(live),
                                                        _ptr(vals), _ptr(rp), _ptr(ci), _ptr(va),
                                                        C.byref(kept)))
        return rp, ci[: kept.value], va[: kept.value]

    def system(self, vals: torch.Tensor, live: Optional[torch.Tensor] = None) -> LinearSystem:
        """The assembled system: on the structural pattern, or with live flags on the
        reference's key set."""
        if live is None:
            rp, ci, va = self.pattern.row_ptr, self.pattern.col_idx, vals[: self.pattern.nnz]
        else:
            rp, ci, va = self.compact(vals, live)
        return LinearSystem(self.n_free, rp, ci, va, self.rhs(), self.dof_eq, self.prescribed)


class Graph:
    """A captured assembly / Newton iteration (cxg_graph_*). Holds the tensors whose
    addresses it captured; update their contents in place between launches."""

    def __init__(self, ctx: Context, h, keep):
        self.ctx, self.h, self._keep = ctx, h, keep
        n = C.c_int64()
        ctx.lib.cxg_graph_info(h, C.byref(n))
        self.nodes = n.value

    def launch(self) -> None:
        self.ctx.check(self.ctx.lib.cxg_graph_launch(self.ctx.h, self.h))

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.ctx.lib.cxg_graph_destroy(h)
            self.h = None


def _only_active(model: Model) -> Model:
    """Drop *MODEL CHANGE, REMOVE elements: they add no stiffness (assembly.cpp:238)."""
    keep = np.asarray(model.active, bool)
    blocks, off = [], 0
    for t, c in model.blocks:
        k = keep[off: off + c.shape[0]]
        if k.any():
            blocks.append((t, c[k]))
        off += c.shape[0]
    return Model(model.coords, blocks, model.E[keep], model.nu[keep], model.spc_node,
                 model.spc_comp, model.spc_value, model.cload_node, model.cload_comp,
                 model.cload_value, rho=model.rho[keep])


def assemble_linear_static(model: Model, device=None) -> LinearSystem:
    """fem::assemble_linear_static (fem/assembly.cpp:225-257) on the GPU."""
    a = Assembler(model, context(device))
    vals = a.new_values()
    live = a.new_live() if a.ctx.tracks_live else None
    a.assemble(vals, live=live)
    return a.system(vals, live)


def _load_free_system(a: Assembler, vals: torch.Tensor, live: torch.Tensor) -> LinearSystem:
    """The compacted system of a matrix that drives no load: assemble_mass and
    assemble_geometric_stiffness discard the eliminated-column side effect."""
    return _key_set_system(a, *a.compact(vals, live))


def _key_set_system(a: Assembler, rp, ci, va) -> LinearSystem:
    rhs = torch.zeros(a.n_free, dtype=torch.float64, device=a.ctx.device)
    return LinearSystem(a.n_free, rp, ci, va, rhs, a.dof_eq, a.prescribed)


def assemble_mass(model: Model, lumped: bool = False, device=None) -> LinearSystem:
    """fem::assemble_mass (fem/assembly.cpp:259-297) with model.rho (*DENSITY):
    consistent element_mass or row-sum element_mass_lumped (fem/element.cpp:786-831).
    *MASS point masses are outside this path."""
    a = Assembler(model, context(device))
    if a.sd_fits():
        nnz = max(a.same_direction()[1].numel(), 1)
        vals = torch.empty(nnz, dtype=torch.float64, device=a.ctx.device)
        live = torch.empty(nnz, dtype=torch.uint8, device=a.ctx.device)
        a.assemble_mass_sd(vals, live, lumped)
        return _key_set_system(a, *a.sd_system(vals, live))
    vals, live = a.new_values(), a.new_live()
    a.assemble_mass(vals, lumped, live=live)
    return _load_free_system(a, vals, live)


def assemble_geometric_stiffness(model: Model, gp_stress, device=None) -> LinearSystem:
    """fem::assemble_geometric_stiffness (fem/assembly.cpp:299-330) with
    element_geometric_stiffness (fem/element.cpp:833-854). gp_stress: per model element
    its (n_gp, 6) Gauss-point stress (recover_gauss_stress's list; removed elements
    empty), or one (n_points, 6) array in model element order."""
    a = Assembler(model, context(device))
    if isinstance(gp_stress, (list, tuple)):
        rows = [np.asarray(g, np.float64).reshape(-1, 6) for g in gp_stress]
        keep = [r for i, r in enumerate(rows)
                if model.active is None or bool(model.active[i])]
        g = np.concatenate(keep) if keep else np.zeros((0, 6))
    else:
        g = np.asarray(gp_stress, np.float64).reshape(-1, 6)
        if model.active is not None and not np.all(model.active):
            ng = np.concatenate([np.full(c.shape[0], GAUSS[t]) for t, c in model.blocks])
            g = g[np.repeat(np.asarray(model.active, bool), ng)]
    gt = torch.from_numpy(np.ascontiguousarray(g))
    if a.sd_fits():
        nnz = max(a.same_direction()[1].numel(), 1)
        vals = torch.empty(nnz, dtype=torch.float64, device=a.ctx.device)
        live = torch.empty(nnz, dtype=torch.uint8, device=a.ctx.device)
        a.assemble_geometric_sd(vals, live, gt)
        return _key_set_system(a, *a.sd_system(vals, live))
    vals, live = a.new_values(), a.new_live()
    a.assemble_geometric(vals, gt, live=live)
    return _load_free_system(a, vals, live)


def internal_force(model: Model, u, device=None) -> torch.Tensor:
    """fem::internal_force (fem/stress.cpp:125-145): f_int (3*n_nodes) = sum_e Ke ue."""
    ctx = context(device)
    mesh = DeviceMesh(model, ctx)
    u = torch.as_tensor(np.asarray(u, np.float64).reshape(-1, 3)).to(ctx.device).contiguous()
    f = torch.empty(3 * model.n_nodes, dtype=torch.float64, device=ctx.device)
    ctx.check(ctx.lib.cxg_internal_force(ctx.h, mesh.h, _ptr(mesh.coords), _ptr(mesh.E),
                                         _ptr(mesh.nu), _ptr(u), _ptr(f)))
    return f


def element_stiffness(etype: int, coords, E: float, nu: float, device=None) -> np.ndarray:
    """fem::element_stiffness (fem/element.cpp:614-634): dense (3n x 3n) Ke, node-major."""
    x = np.asarray(coords, np.float64).reshape(-1, 3)
    n = NODES[etype]
    if x.shape[0] != n:
        raise ValueError("coords do not match the element type")
    m = Model(x, [(etype, np.arange(n, dtype=np.int32)[None, :])], E, nu)
    sys = assemble_linear_static(m, device)
    K = np.zeros((3 * n, 3 * n))
    rows, cols, vals = sys.coo()
    K[rows, cols] = vals
    return K


class MaterialPoints:
    """Per-Gauss-point J2 history on the device (fem/assembly.hpp:76-93): committed
    and trial planes, each (13, n_points) float64 SoA."""

    def __init__(self, n_points: int, device=None, committed: Optional[np.ndarray] = None):
        dev = _dev(device)
        self.n_points = n_points
        if committed is None:
            self.committed = torch.zeros((_lib.STATE_COMPONENTS, n_points), dtype=torch.float64, device=dev)
        else:
            self.committed = torch.from_numpy(np.ascontiguousarray(committed, np.float64)).to(dev)
        self.trial = self.committed.clone()

    def commit(self):
        """MaterialPoints::commit (fem/assembly.hpp:84-87)."""
        ctx = context(self.committed.device)
        ctx.check(ctx.lib.cxg_commit(ctx.h, self.n_points, _ptr(self.committed), _ptr(self.trial)))


def evaluate_material_points(strain, mats=(), mat_id=None, E=None, nu=None, committed=None,
                             device=None):
    """MaterialModel::evaluate (fem/material_model.hpp:77) at n independent points on the
    device (cxg_material_evaluate): J2PlasticMaterial::evaluate (plasticity.cpp:90-236)
    where mat_id[p] >= 0 names one of `mats` (J2Params), ElasticIsoMaterial with E[p],
    nu[p] where it is -1. strain (n, 6), committed (13, n). Returns host arrays
    (stress (n, 6), tangent (n, 6, 6), trial (13, n)); trial holds the committed history
    at elastic points."""
    dev = _dev(device)
    ctx = context(dev)
    eps = torch.from_numpy(np.ascontiguousarray(strain, np.float64)).to(dev)
    n = eps.shape[0]
    hc = torch.zeros((_lib.STATE_COMPONENTS, n), dtype=torch.float64, device=dev) \
        if committed is None else torch.from_numpy(np.ascontiguousarray(committed, np.float64)).to(dev)
    trial = hc.clone()
    table = (_lib.J2Material * max(len(mats), 1))()
    for i, m in enumerate(mats):
        table[i] = m.to_c()
    mid = None if mat_id is None else torch.from_numpy(np.ascontiguousarray(mat_id, np.int32)).to(dev)
    Ed = None if E is None else torch.from_numpy(np.broadcast_to(np.asarray(E, np.float64), (n,)).copy()).to(dev)
    nud = None if nu is None else torch.from_numpy(np.broadcast_to(np.asarray(nu, np.float64), (n,)).copy()).to(dev)
    sig = torch.empty((n, 6), dtype=torch.float64, device=dev)
    D = torch.empty((n, 6, 6), dtype=torch.float64, device=dev)
    ctx.check(ctx.lib.cxg_material_evaluate(ctx.h, n, _ptr(mid), table, len(mats), _ptr(Ed),
                                            _ptr(nud), _ptr(eps), _ptr(hc), _ptr(trial),
                                            _ptr(sig), _ptr(D)))
    return sig.cpu().numpy(), D.cpu().numpy(), trial.cpu().numpy()


def make_material_points(model: Model, device=None) -> MaterialPoints:
    """fem::make_material_points (fem/assembly.cpp:357-383): virgin history."""
    return MaterialPoints(model.n_points, device)


class TangentAssembler(Assembler):
    """assemble_material_tangent with the mesh and pattern kept resident."""

    def __init__(self, model: Model, ctx: Optional[Context] = None):
        if model.active is not None and not np.all(model.active):
            raise NotImplementedError("*MODEL CHANGE on the material-point path")
        super().__init__(model, ctx)
        dev = self.ctx.device
        self.mats = (_lib.J2Material * max(len(model.j2), 1))()
        for i, p in enumerate(model.j2):
            self.mats[i] = p.to_c()
        self.n_mats = len(model.j2)
        jid = model.j2_id if model.j2_id is not None else np.full(model.n_elem, -1, np.int32)
        self.mat_id = torch.from_numpy(jid).to(dev)

    def assemble_tangent(self, u: torch.Tensor, mp: MaterialPoints, vals: torch.Tensor,
                         f_int: torch.Tensor, live: Optional[torch.Tensor] = None) -> None:
        m = self.mesh
        self.ctx.check(self.ctx.lib.cxg_assemble_j2(
            self.ctx.h, self.pattern.h, _ptr(m.coords), _ptr(m.E), _ptr(m.nu), _ptr(self.mat_id),
            self.mats, self.n_mats, _ptr(u), _ptr(mp.committed), _ptr(mp.trial), _ptr(vals),
            _ptr(f_int), _ptr(live)))

    def capture(self, u: torch.Tensor, mp: MaterialPoints, vals: torch.Tensor,
                f_int: torch.Tensor, f_ext: Optional[torch.Tensor] = None,
                rhs: Optional[torch.Tensor] = None) -> Graph:
        """One Newton iteration (return map + K_t + f_int, and rhs = f_ext - f_int on
        the free DOFs when rhs is given) captured as a HIP graph."""
        m = self.mesh
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.cxg_graph_capture_j2(
            self.ctx.h, self.pattern.h, _ptr(m.coords), _ptr(m.E), _ptr(m.nu), _ptr(self.mat_id),
            self.mats, self.n_mats, _ptr(u), _ptr(mp.committed), _ptr(mp.trial), _ptr(vals),
            _ptr(f_int), _ptr(f_ext), _ptr(rhs), C.byref(h)))
        return Graph(self.ctx, h, (self, u, mp.committed, mp.trial, vals, f_int, f_ext, rhs))


def assemble_material_tangent(model: Model, u, mp: MaterialPoints, device=None):
    """fem::assemble_material_tangent (fem/assembly.cpp:385-423): returns
    (LinearSystem with K_t, f_int over the full DOF set); advances mp's trial history."""
    a = TangentAssembler(model, context(device))
    u = torch.as_tensor(np.asarray(u, np.float64).reshape(-1, 3) if not torch.is_tensor(u) else u)
    u = u.to(a.ctx.device, torch.float64).contiguous()
    vals = a.new_values()
    live = a.new_live() if a.ctx.tracks_live else None
    f_int = torch.empty(3 * model.n_nodes, dtype=torch.float64, device=a.ctx.device)
    a.assemble_tangent(u, mp, vals, f_int, live)
    return a.system(vals, live), f_int
