"""Write tests/golden/golden.npz: small fixtures computed by the oracle (oracle/, the CPU
restatement of the reference's algorithms). Run from the repository root:

    python tests/golden/make_golden.py

Contents:
* ``c1_*``: links, q, q >= 1/2 and Bouzidi weights of the 64x32x32 sphere channel
  (sphere r = 6 at (20, 16, 16), analytic SDF), from the integer link builder and the
  FP32 q search (sdf_tracing.py:20-259).
* ``c1_s{1,10,100}_{rho,u,v,w}``: the z = 16 plane of the macros after 1, 10 and 100
  steps of that channel (BGK tau = 0.6, Zou/He inlet u = 0.05 at x = 0, convective
  outlet, walls, Bouzidi links), FP32 oracle step (lbm_cuda_ext.py:445-517);
  ``c1_s{...}_mass``: the FP64 sum of rho over the whole grid.
* ``mg_*``: the 32x16x16 sphere-conduction problem (k_s/k_f = 100, walls at
  T = 0, omega = 0.6, V(2,2), 2 levels): an input ``mg_r``, one V-cycle of it in FP32
  (``mg_z32``) and FP64 (``mg_z64``), and the FP64 flexible-CG solution ``mg_x64`` to a
  relative residual of 1e-10.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")
STEPS = (1, 10, 100)
PLANE = 16
MG_SHAPE = (32, 16, 16)
MG_OMEGA = 0.6


def config1_bcs(nx):
    return [dict(kind="zou_he", axis=0, plane=0, interior=1, bc_type=0, u_in=0.05, rho_out=1.0, sign=1),
            dict(kind="convective", axis=0, plane=nx - 1, interior=nx - 2, unknown_cax=-1, rho_out=1.0)]


def make_links():
    from fluxcore3d_amd import geometry
    from oracle import lbm as olbm
    phi, solid, solid_flow = geometry.sphere_channel()
    L = olbm.ibb_from_sdf(solid, phi)
    return {f"c1_{k}": np.asarray(L[k]) for k in ("fluid_flat", "dir_i", "ff_flat", "ff_valid", "q", "q_ge_half",
                                                  "w1", "w2")}


def make_steps():
    from fluxcore3d_amd import geometry
    from oracle import lbm as olbm
    phi, solid, solid_flow = geometry.sphere_channel()
    nx, ny, nz = solid.shape
    links = olbm.ibb_from_sdf(solid, phi)
    f = np.broadcast_to(olbm.W32[:, None], (19, nx * ny * nz)).copy()
    out = {}
    for step in range(1, max(STEPS) + 1):
        f, rho, u, v, w = olbm.step(f, solid_flow, nx, ny, nz, (0, 0, 0), "bgk", 1.0 / 0.6, 1.0, None, 0.0, links,
                                    None, config1_bcs(nx))
        if step in STEPS:
            for name, a in (("rho", rho), ("u", u), ("v", v), ("w", w)):
                out[f"c1_s{step}_{name}"] = a.reshape(nx, ny, nz)[:, :, PLANE].astype(np.float32)
            out[f"c1_s{step}_mass"] = np.float64(rho.astype(np.float64).sum())
    return out


def mg_case():
    from tests import thermal_cases
    return thermal_cases.sphere_conduction(*MG_SHAPE, dt=1.0e6)


def make_mg():
    from oracle import mg
    c = mg_case()
    out = {}
    r = (np.random.default_rng(5).standard_normal(MG_SHAPE) * c["mask"]).astype(np.float32)
    out["mg_r"] = r
    for dtype, key in ((np.float32, "mg_z32"), (np.float64, "mg_z64")):
        h = mg.Hierarchy(*[c[k].astype(dtype) for k in ("k", "rcp", "u", "v", "w")], c["mask"], c["ax"], c["sign"],
                         c["s"], omega=MG_OMEGA)
        h.set_dt(c["dt"])
        out[key] = np.asarray(h.precond(r.astype(dtype)), dtype=dtype)
    h = mg.Hierarchy(c["k"], c["rcp"], c["u"], c["v"], c["w"], c["mask"], c["ax"], c["sign"], c["s"], omega=MG_OMEGA)
    h.set_dt(c["dt"])
    x, _ = mg.fcg_mg_fp64(h, c["b"], tol=1e-10, max_it=300)
    out["mg_x64"] = np.asarray(x, dtype=np.float64)
    return out


def make():
    out = {}
    out.update(make_links())
    out.update(make_steps())
    out.update(make_mg())
    return out


if __name__ == "__main__":
    data = make()
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT}: {len(data)} arrays, {os.path.getsize(OUT)} bytes")
