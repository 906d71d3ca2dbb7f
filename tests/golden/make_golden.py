"""Write the golden fixtures of the assembly path from the reference core.

Runs only where the reference tree is mounted: it needs oracle/_ref/libcxpp_ref.so
(built by `make -C oracle ref` from the reference's own calculixpp_core sources).
Every array in the fixtures is an output of a reference function called through
oracle/ref_shim.cpp; inputs are seeded. The fixtures are data (inputs + expected
outputs), committed so that tests pin the oracle without the reference present.

    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py recover    # one fixture group (or mass_geometric)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from calculixpp_amd import _lib, meshgen  # noqa: E402
from calculixpp_amd.model import J2Params, Model  # noqa: E402
from oracle.oracle import Oracle, aos_state  # noqa: E402

FAMILIES = {"C3D8": _lib.C3D8, "C3D20": _lib.C3D20, "C3D20R": _lib.C3D20R, "C3D6": _lib.C3D6}


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


def structural_keys(model, dof_eq):
    """Free DOF pairs of nodes that share an element (the structural pattern)."""
    keys = set()
    for _, conn in model.blocks:
        for e in conn:
            dofs = [dof_eq[3 * n + c] for n in e for c in range(3)]
            dofs = [d for d in dofs if d >= 0]
            for a in dofs:
                for b in dofs:
                    keys.add((a, b))
    return keys


def system_arrays(sys_, prefix=""):
    rp, ci, v = sys_.csr()
    return {prefix + "row_ptr": rp, prefix + "col_idx": ci, prefix + "vals": v,
            prefix + "rhs": sys_.rhs, prefix + "dof_eq": sys_.dof_eq}


def recover(R):
    """6. Field recovery: recover_fields and recover_gauss_stress on the 2x2x2 blocks,
    the loaded cantilever, and the material-point form on a J2 block."""
    rec = {}
    for name, t in FAMILIES.items():
        m = meshgen.block(2, t, seed_geo=21, seed_E=22)
        u = meshgen.smooth_field(m.coords, 41)
        o = R.recover_fields(m, u, gauss_stress=True)
        rec[name + "_u"] = u
        for k in ("strain", "stress", "reaction", "gauss_stress"):
            rec[name + "_" + k] = o[k]
    m = meshgen.cantilever()
    u = meshgen.smooth_field(m.coords, 42)
    o = R.recover_fields(m, u)
    rec["cantilever_u"] = u
    for k in ("strain", "stress", "reaction"):
        rec["cantilever_" + k] = o[k]
    m = meshgen.j2_block(3, 2, 2, (1.0, 0.8, 0.6), family=_lib.C3D20, elastic_every=5,
                         jitter_frac=0.1)
    hist = meshgen.seeded_history(m.n_points, 43)
    u = meshgen.smooth_field(m.coords, 44, 1.5e-4)
    o = R.recover_fields(m, u, aos_state(hist))
    rec["j2_u"], rec["j2_history"] = u, hist
    for k in ("strain", "stress", "reaction"):
        rec["j2_" + k] = o[k]
    save("recover", **rec)


def mass_geometric(R):
    """7. assemble_mass (consistent, lumped) and assemble_geometric_stiffness on the
    2x2x2 blocks with per-element densities and seeded Gauss-point stresses."""
    out = {}
    rng = np.random.default_rng(45)
    for name, t in FAMILIES.items():
        b = meshgen.block(2, t, seed_geo=21, seed_E=22)
        rho = rng.uniform(1e-9, 8e-9, b.n_elem)
        gs = rng.normal(scale=100.0, size=(b.n_points, 6))
        m = Model(b.coords, b.blocks, b.E, b.nu, rho=rho)
        out[name + "_rho"], out[name + "_gp_stress"] = rho, gs
        for key, s in (("mass", R.assemble_mass(m)), ("lumped", R.assemble_mass(m, True)),
                       ("geo", R.assemble_geometric_stiffness(m, gs))):
            out.update(system_arrays(s, f"{name}_{key}_"))
    save("mass_geometric", **out)


def main():
    R = Oracle("reference")
    if sys.argv[1:] == ["recover"]:
        recover(R)
        return
    if sys.argv[1:] == ["mass_geometric"]:
        mass_geometric(R)
        return
    # 1. config 1 cantilever: K_ff, rhs; the reference key set equals the structural set
    m = meshgen.cantilever()
    s = R.assemble_linear_static(m)
    assert set(zip(s.rows.tolist(), s.cols.tolist())) == structural_keys(m, s.dof_eq)
    save("cantilever_c3d8", **system_arrays(s))

    # 2. single elements, jittered, one per family
    rng = np.random.default_rng(20)
    ele = {}
    for name, t in FAMILIES.items():
        coords, conn = meshgen.hex_grid(1, 1, 1, (1.0, 1.0, 1.0), t)
        x = coords[conn[0]] + rng.uniform(-0.08, 0.08, size=(conn.shape[1], 3))
        E, nu = rng.uniform(1e5, 3e5), rng.uniform(0.2, 0.4)
        ele[name + "_coords"] = x
        ele[name + "_Enu"] = np.array([E, nu])
        ele[name + "_Ke"] = R.element_stiffness(t, x, E, nu)
        ue = rng.normal(scale=1e-3, size=x.shape)
        Ke2, fe, _ = R.element_tangent_force(t, x, ue, E, nu)
        ele[name + "_ue"] = ue
        ele[name + "_fe"] = fe
    save("elements", **ele)

    # 3. 2x2x2 jittered blocks per family: K, f_int
    blk = {}
    for name, t in FAMILIES.items():
        m = meshgen.block(2, t, seed_geo=21, seed_E=22)
        s = R.assemble_linear_static(m)
        assert set(zip(s.rows.tolist(), s.cols.tolist())) == structural_keys(m, s.dof_eq)
        u = meshgen.smooth_field(m.coords, 23)
        blk.update(system_arrays(s, name + "_"))
        blk[name + "_u"] = u
        blk[name + "_fint"] = R.internal_force(m, u)
    save("blocks_2x2x2", **blk)

    # 4. J2 material points: isotropic (config 3 table), kinematic, combined
    mats = {"iso": J2Params(210000.0, 0.3, [0.0, 1.0], [250.0, 2250.0]),
            "kin": J2Params(200000.0, 0.29, [0.0, 0.5], [300.0, 800.0], _lib.HARD_KINEMATIC),
            "comb": J2Params(70000.0, 0.33, [0.0, 0.01, 0.1], [100.0, 150.0, 180.0],
                             _lib.HARD_COMBINED)}
    pts = {}
    n = 48
    for key, p in mats.items():
        hist = meshgen.seeded_history(n, seed=30)
        strain = rng.normal(scale=2.5e-3, size=(n, 6))
        st0 = aos_state(hist)
        sig, D, st = np.zeros((n, 6)), np.zeros((n, 6, 6)), np.zeros_like(st0)
        for i in range(n):
            sig[i], D[i], st[i] = R.j2_evaluate(p, strain[i], st0[i])
        pts[key + "_strain"], pts[key + "_state_in"] = strain, st0
        pts[key + "_stress"], pts[key + "_D"], pts[key + "_state_out"] = sig, D, st
        pts[key + "_params"] = np.array([p.E, p.nu, p.hardening, p.kinematic_modulus])
        pts[key + "_table"] = np.array([p.eqplastic, p.yield_])
    save("j2_points", **pts)

    # 5. J2 tangent assembly on a 3x3x2 C3D8 mesh (config 3 material, seeded history)
    m = meshgen.j2_block(3, 3, 2, (1.5, 1.5, 1.0))
    hist = meshgen.seeded_history(m.n_points, 3)
    u = meshgen.smooth_field(m.coords, 7, 2e-3)
    s, fi, st = R.assemble_material_tangent(m, u, aos_state(hist))
    arr = system_arrays(s)
    save("j2_mesh_c3d8", u=u, history=hist, f_int=fi, state_out=st, **arr)

    recover(R)
    mass_geometric(R)


if __name__ == "__main__":
    main()
