"""Seeded structured meshes and the workloads of the assembly path.

Node numbering is x fastest, then y, then z (config 1's node id 1+i+21(j+5k)).
Element node order is CalculiX's (corners 1-8, then midside nodes 9-20 on edges
1-2, 2-3, 3-4, 4-1, 5-6, 6-7, 7-8, 8-5, 1-5, 2-6, 3-7, 4-8; src/fem/element.cpp:
160-185). C3D6 wedges split each hex cell into (1,2,3 | 5,6,7) and (1,3,4 | 5,7,8).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .model import J2Params, Model


def hex_grid(nx: int, ny: int, nz: int, lengths=(1.0, 1.0, 1.0), family: int = _lib.C3D8):
    """(coords (n,3), conn (ne, nn) int32) of a structured box [0,Lx]x[0,Ly]x[0,Lz]."""
    lx, ly, lz = lengths
    if family in (_lib.C3D8, _lib.C3D6):
        xs = np.linspace(0.0, lx, nx + 1)
        ys = np.linspace(0.0, ly, ny + 1)
        zs = np.linspace(0.0, lz, nz + 1)
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        i, j, k = i.ravel(), j.ravel(), k.ravel()

        def nid(a, b, c):
            return (a + (nx + 1) * (b + (ny + 1) * c)).astype(np.int32)
        c = [nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
             nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)]
        hexes = np.stack(c, axis=1)
        if family == _lib.C3D8:
            return coords, hexes
        w1 = hexes[:, [0, 1, 2, 4, 5, 6]]
        w2 = hexes[:, [0, 2, 3, 4, 6, 7]]
        wed = np.empty((2 * len(hexes), 6), np.int32)
        wed[0::2] = w1
        wed[1::2] = w2
        return coords, wed
    if family in (_lib.C3D20, _lib.C3D20R):
        LX, LY, LZ = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1
        K, J, I = np.meshgrid(np.arange(LZ), np.arange(LY), np.arange(LX), indexing="ij")
        odd = (I % 2) + (J % 2) + (K % 2)
        keep = odd <= 1
        index = np.full(keep.shape, -1, np.int64)
        index[keep] = np.arange(int(keep.sum()))
        coords = np.stack([I[keep] * (lx / (2 * nx)), J[keep] * (ly / (2 * ny)),
                           K[keep] * (lz / (2 * nz))], axis=1).astype(np.float64)
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        i, j, k = 2 * i.ravel(), 2 * j.ravel(), 2 * k.ravel()

        def at(a, b, c):
            return index[c, b, a].astype(np.int32)
        cols = [at(i, j, k), at(i + 2, j, k), at(i + 2, j + 2, k), at(i, j + 2, k),
                at(i, j, k + 2), at(i + 2, j, k + 2), at(i + 2, j + 2, k + 2), at(i, j + 2, k + 2),
                at(i + 1, j, k), at(i + 2, j + 1, k), at(i + 1, j + 2, k), at(i, j + 1, k),
                at(i + 1, j, k + 2), at(i + 2, j + 1, k + 2), at(i + 1, j + 2, k + 2), at(i, j + 1, k + 2),
                at(i, j, k + 1), at(i + 2, j, k + 1), at(i + 2, j + 2, k + 1), at(i, j + 2, k + 1)]
        return coords, np.stack(cols, axis=1)
    raise ValueError(f"family {family} outside the GPU path")


def jitter(coords: np.ndarray, h, frac: float, seed: int, lengths) -> np.ndarray:
    """Move every node strictly inside the box by U[-frac*h, frac*h] per component."""
    rng = np.random.default_rng(seed)
    h = np.broadcast_to(np.asarray(h, np.float64), (3,))
    L = np.asarray(lengths, np.float64)
    tol = 1e-12 * L.max()
    interior = np.all((coords > tol) & (coords < L - tol), axis=1)
    d = rng.uniform(-1.0, 1.0, size=coords.shape) * (frac * h)
    out = coords.copy()
    out[interior] += d[interior]
    return out


def smooth_field(coords: np.ndarray, seed: int, amplitude: float = 1e-3, terms: int = 4) -> np.ndarray:
    """u_c(x) = sum_m a_m sin(pi k_m . x + phi_m) per component, scaled to max|u| = amplitude."""
    rng = np.random.default_rng(seed)
    u = np.zeros_like(coords)
    for c in range(3):
        for _ in range(terms):
            k = rng.integers(1, 4, size=3).astype(np.float64)
            phi = rng.uniform(0.0, 2.0 * np.pi)
            a = rng.uniform(-1.0, 1.0)
            u[:, c] += a * np.sin(np.pi * coords @ k + phi)
    m = np.abs(u).max()
    return u * (amplitude / m if m > 0 else 1.0)


# ---- the five workloads -----------------------------------------------------------

def cantilever(nx: int = 20, ny: int = 4, nz: int = 4, lengths=(1.0, 0.2, 0.2),
               E: float = 210000.0, nu: float = 0.3, traction: float = -1.0) -> Model:
    """Config 1: C3D8 prism clamped at x=0 with a uniform z traction on x=L entered as
    consistent nodal *CLOADs (corner 1/4, edge 1/2, interior 1 of the cell share)."""
    coords, conn = hex_grid(nx, ny, nz, lengths, _lib.C3D8)
    L = lengths[0]
    x0 = np.nonzero(np.abs(coords[:, 0]) < 1e-12)[0]
    spc_node = np.repeat(x0, 3).astype(np.int32)
    spc_comp = np.tile(np.array([1, 2, 3], np.int32), len(x0))
    tip = np.nonzero(np.abs(coords[:, 0] - L) < 1e-12)[0]
    cell = (lengths[1] / ny) * (lengths[2] / nz) * traction
    yb = np.isclose(coords[tip, 1], 0) | np.isclose(coords[tip, 1], lengths[1])
    zb = np.isclose(coords[tip, 2], 0) | np.isclose(coords[tip, 2], lengths[2])
    share = np.where(yb & zb, 0.25, np.where(yb | zb, 0.5, 1.0))
    return Model(coords, [(_lib.C3D8, conn)], E, nu, spc_node=spc_node, spc_comp=spc_comp,
                 spc_value=np.zeros(len(spc_node)), cload_node=tip.astype(np.int32),
                 cload_comp=np.full(len(tip), 3, np.int32), cload_value=share * cell)


def block(n: int = 64, family: int = _lib.C3D20, seed_geo: int = 0, seed_E: int = 1,
          jitter_frac: float = 0.1, ny: int | None = None, nz: int | None = None) -> Model:
    """Config 2: [0,1]^3 block, interior nodes jittered +-10% h, E_e ~ U[150000, 250000]."""
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    L = (1.0, ny / n, nz / n)
    coords, conn = hex_grid(n, ny, nz, L, family)
    coords = jitter(coords, 1.0 / n, jitter_frac, seed_geo, L)
    E = np.random.default_rng(seed_E).uniform(150000.0, 250000.0, size=conn.shape[0])
    return Model(coords, [(family, conn)], E, 0.3)


def j2_block(nx: int = 128, ny: int = 128, nz: int = 64, lengths=(2.0, 2.0, 1.0),
             family: int = _lib.C3D8, elastic_every: int = 0, jitter_frac: float = 0.0,
             seed_geo: int = 9):
    """Config 3 mesh + material: C3D8, E=210000, nu=0.3, sigma_y0 = 250, H = 2000
    (table {(0, 250), (1, 2250)}). Other families on the same grid for parity runs;
    elastic_every = k > 0 leaves every k-th element isotropic elastic (no J2);
    jitter_frac > 0 moves interior nodes by up to that fraction of the x spacing."""
    coords, conn = hex_grid(nx, ny, nz, lengths, family)
    if jitter_frac > 0.0:
        coords = jitter(coords, lengths[0] / nx, jitter_frac, seed_geo, lengths)
    mat = J2Params(210000.0, 0.3, [0.0, 1.0], [250.0, 2250.0])
    jid = np.zeros(conn.shape[0], np.int32)
    if elastic_every > 0:
        jid[::elastic_every] = -1
    m = Model(coords, [(family, conn)], 210000.0, 0.3, j2_id=jid, j2=[mat])
    return m


def seeded_history(n_points: int, seed: int = 3) -> np.ndarray:
    """Committed J2 history (13, n_points): deviatoric plastic strain |.| <= 1e-3
    (engineering shear), ep ~ U[0, 2e-3], deviatoric back stress |.| <= 10."""
    rng = np.random.default_rng(seed)
    h = np.zeros((13, n_points))
    ep = rng.uniform(-1e-3, 1e-3, size=(6, n_points))
    ep[:3] -= ep[:3].mean(axis=0)
    h[0:6] = ep
    h[6] = rng.uniform(0.0, 2e-3, size=n_points)
    al = rng.uniform(-10.0, 10.0, size=(6, n_points))
    al[:3] -= al[:3].mean(axis=0)
    h[7:13] = al
    return h


def adjoint_block(nx: int = 96, ny: int = 48, nz: int = 48, lengths=(2.0, 1.0, 1.0),
                  E0: float = 210000.0, penal: float = 3.0):
    """Config 5: C3D8 mesh jittered +-5% h (seed 5), SIMP density rho ~ U[0.05, 1]
    (seed 6), E = E0 rho^3; returns (model, rho)."""
    coords, conn = hex_grid(nx, ny, nz, lengths, _lib.C3D8)
    coords = jitter(coords, lengths[0] / nx, 0.05, 5, lengths)
    rho = np.random.default_rng(6).uniform(0.05, 1.0, size=conn.shape[0])
    return Model(coords, [(_lib.C3D8, conn)], E0 * rho ** penal, 0.3), rho


def mixed_box(n_elem: int, seed: int, jitter_frac: float = 0.1):
    """A structured box of about n_elem elements whose cells are drawn uniformly as one
    C3D8, one C3D20 or two C3D6 (a hex split along its 0-2 / 4-6 diagonal plane).
    Nodes live on the quadratic lattice; unused lattice nodes are dropped. Interior
    nodes are jittered by up to jitter_frac of the spacing. Returns (coords, blocks)
    with blocks in the order C3D8, C3D20, C3D6 (empty runs left out)."""
    rng = np.random.default_rng(seed)
    cells = max(1, int(round(n_elem * 0.75)))  # 4/3 elements per cell on average
    nx = max(1, int(round(cells ** (1.0 / 3.0))))
    ny = max(1, int(round((cells / nx) ** 0.5)))
    nz = max(1, int(round(cells / (nx * ny))))
    h = 1.0 / max(nx, ny, nz)
    L = (nx * h, ny * h, nz * h)
    coords, q = hex_grid(nx, ny, nz, L, _lib.C3D20)
    kind = rng.integers(0, 3, size=q.shape[0])
    c8 = q[kind == 0, :8]
    c20 = q[kind == 1]
    hx = q[kind == 2, :8]
    c6 = np.empty((2 * len(hx), 6), np.int32)
    c6[0::2] = hx[:, [0, 1, 2, 4, 5, 6]]
    c6[1::2] = hx[:, [0, 2, 3, 4, 6, 7]]
    used = np.zeros(len(coords), bool)
    for c in (c8, c20, c6):
        used[c.ravel()] = True
    renum = np.full(len(coords), -1, np.int64)
    renum[used] = np.arange(int(used.sum()))
    coords = jitter(coords[used], h, jitter_frac, seed + 1, L)
    blocks = [(t, renum[c].astype(np.int32)) for t, c in
              ((_lib.C3D8, c8), (_lib.C3D20, c20), (_lib.C3D6, c6)) if len(c)]
    return coords, blocks


J2_TABLE = ([0.0, 1.0], [250.0, 2250.0])


def batch_meshes(n_meshes: int = 512, seed: int = 4, e_min: int = 100, e_max: int = 50000,
                 n_cases: int = 4):
    """Config 4: n_meshes independent mixed-family boxes with element counts log-uniform
    in [e_min, e_max] (seed 4), per-mesh E ~ U[70000, 210000] and nu ~ U[0.25, 0.35];
    half of the meshes (a seeded draw) take a J2 step with the table {(0,250),(1,2250)}
    and seeded history. Each mesh carries n_cases smooth displacement fields (max |u|
    1e-3 for linear meshes, 2e-4 for J2 meshes). Returns a list of dicts with keys
    model, j2 (bool), u (n_cases, n_nodes, 3) and, for J2 meshes, history (13, n_points)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(e_min), np.log(e_max), size=n_meshes)).astype(np.int64)
    E = rng.uniform(70000.0, 210000.0, size=n_meshes)
    nu = rng.uniform(0.25, 0.35, size=n_meshes)
    j2 = np.zeros(n_meshes, bool)
    j2[rng.choice(n_meshes, n_meshes // 2, replace=False)] = True
    out = []
    for i in range(n_meshes):
        ms = 1000 * (seed + 1) + 10 * i
        coords, blocks = mixed_box(int(sizes[i]), ms)
        n_el = sum(c.shape[0] for _, c in blocks)
        kw = {}
        if j2[i]:
            kw = dict(j2_id=np.zeros(n_el, np.int32),
                      j2=[J2Params(float(E[i]), float(nu[i]), *J2_TABLE)])
        m = Model(coords, blocks, float(E[i]), float(nu[i]), **kw)
        amp = 2e-4 if j2[i] else 1e-3
        u = np.stack([smooth_field(coords, ms + 2 + c, amp) for c in range(n_cases)])
        rec = {"model": m, "j2": bool(j2[i]), "u": u}
        if j2[i]:
            rec["history"] = seeded_history(m.n_points, ms + 7)
        out.append(rec)
    return out
