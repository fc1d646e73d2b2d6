"""Writes tests/golden/gyroid_heat_fp64.npz: the FP64 oracle of config 4's conjugate
heat leg (tests/cases.gyroid_heat_params) on one period of the gyroid (48^3,
periodic), which tiles the 96^3 crop and the 384^3 domain exactly: the Nusselt
number and layer means after 1000 and 5000 steps, and the fields u, v, w, T after
5000 steps at every second node along each axis.

Run from the repository root: python tests/golden/make_gyroid_heat_fixture.py
(about four minutes on eight cores)."""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))

from oracle import d3q7 as od  # noqa: E402
from oracle import lbm as olbm  # noqa: E402
from tests import cases  # noqa: E402


def main():
    p = cases.gyroid_heat_params()
    c = cases.gyroid_box(n=48, period=48.0)
    n = 48
    N = n ** 3
    solid = c["solid"].reshape(-1)
    fl = np.where(solid, od.CONJ, od.FLUID).astype(np.uint8)
    n_s = int(solid.sum())
    Q = np.where(solid, p["q_solid"], -p["q_solid"] * n_s / (N - n_s))
    kf = p["kappa_f"]
    om_f, om_s = 1.0 / (0.5 + 4 * kf), 1.0 / (0.5 + 4 * p["k_ratio"] * kf)
    f = np.broadcast_to(olbm.W32[:, None].astype(np.float64), (19, N)).copy()
    z = np.zeros(N)
    g = od.equilibrium(z, z, z, z)
    S = p["S"].astype(np.float64)
    out = {}
    t0 = time.time()
    for s in range(1, p["steps"] + 1):
        f, rho, u, v, w = olbm.step(f, c["solid"], n, n, n, (1, 1, 1), "mrt", p["omega"], 1.0, S, 0.0, c["links"],
                                    force=p["force"])
        g, T = od.step(g, fl, u, v, w, z, (n, n, n), periodic=(1, 1, 1), omega=om_f, omega_s=om_s, Q=Q)
        if s in (1000, p["steps"]):
            nu = od.nusselt(g, T, fl, (n, n, n), kf)
            for k, val in nu.items():
                out[f"{k}_{s}"] = np.float64(val)
            print(f"step {s}: {nu} ({time.time() - t0:.0f} s)", flush=True)
    sub = (slice(None, None, 2),) * 3
    for name, a in (("u", u), ("v", v), ("w", w), ("T", T)):
        out[name] = a.reshape(n, n, n)[sub].astype(np.float32)
    out["steps"] = np.int64(p["steps"])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gyroid_heat_fp64.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
