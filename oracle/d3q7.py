"""D3Q7 temperature transport and the coupled flow/solid heat-transfer loop — TEST
INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference carries temperature with an implicit finite-volume upwind operator
(solver/nlevel_mg_cuda.py:418-507, LBM_CHT.py:2208-2359) after the flow has been
frozen (UI_components/cht_worker.py:23-165). The coupled loop restated here has
no counterpart in the reference: the fluid temperature advances every flow step
on the D3Q7 lattice (orc_d3q7_step in lbm_oracle.c: textbook BGK model,
Chopard, Falcone & Latt 2009), and every ``couple_every`` steps the conducting
solid takes one implicit conduction step over the elapsed time with the energy
the fluid handed to it as a source, solved here by a sparse direct FP64 solve.

Pins (tests/test_oracle_d3q7.py): the analytic 1-D advection-diffusion of a
Gaussian pulse (with the model's known diffusivity (1/omega - 1/2)(1/4 - u^2)
along the flow), the linear conduction profile between two solids held at
different temperatures, exact energy conservation with walls, solids and
periodic faces, the solid solve against the conduction balance it encodes, the
conjugate (carried-solid) steady state against the harmonic-mean finite-volume
solution, exact tiling of periodic domains, and the Nusselt diagnostic against
the steady periodic slab solved by finite volumes.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, Optional

import numpy as np

from . import lbm as olbm

CX7 = np.array([0, 1, -1, 0, 0, 0, 0])
CY7 = np.array([0, 0, 0, 1, -1, 0, 0])
CZ7 = np.array([0, 0, 0, 0, 0, 1, -1])
W7 = np.array([0.25] + [0.125] * 6)

FLUID, SOLID, WALL, CONJ, INLET = 0, 1, 2, 3, 16


def flags(conducting: np.ndarray, wall: Optional[np.ndarray] = None, inlet: Optional[np.ndarray] = None) -> np.ndarray:
    """uint8 node kinds: 1 conducting solid, 2 adiabatic wall (where not conducting),
    0 fluid, bit 16 on fluid inlet nodes."""
    fl = np.zeros(conducting.shape, np.uint8)
    if wall is not None:
        fl[np.asarray(wall, bool)] = WALL
    fl[np.asarray(conducting, bool)] = SOLID
    if inlet is not None:
        fl[np.asarray(inlet, bool) & (fl == FLUID)] |= INLET
    return fl


def equilibrium(T, u, v, w, dtype=np.float64) -> np.ndarray:
    """g_i = w_i T (1 + 4 c_i.u), shape (7, N)."""
    T, u, v, w = (np.asarray(a, dtype).reshape(-1) for a in (T, u, v, w))
    out = np.empty((7, T.size), dtype)
    for q in range(7):
        cu = dtype(CX7[q]) * u + dtype(CY7[q]) * v + dtype(CZ7[q]) * w
        out[q] = dtype(W7[q]) * T * (dtype(1) + dtype(4) * cu)
    return out


def step(g, fl, u, v, w, T_solid, shape, periodic=(0, 0, 0), omega=1.0, T_in=0.0, Q=None, acc=None, omega_s=None):
    """One D3Q7 step (orc_d3q7_step). Returns (g_next, T); ``acc`` (solid energy
    accumulator, same dtype) is updated in place; ``omega_s`` relaxes kind-3 nodes
    (default omega)."""
    dtype = g.dtype.type
    lib = olbm._lib(dtype)
    nx, ny, nz = shape
    N = nx * ny * nz
    gin = np.ascontiguousarray(g.reshape(7, N))
    gout = np.zeros_like(gin)
    T = np.zeros(N, dtype)
    arr = [np.ascontiguousarray(np.asarray(a, dtype).reshape(-1)) for a in (u, v, w, T_solid)]
    fl = np.ascontiguousarray(fl.reshape(-1), np.uint8)
    Qa = None if Q is None else np.ascontiguousarray(np.asarray(Q, dtype).reshape(-1))
    if acc is not None:
        assert acc.dtype == gin.dtype and acc.flags.c_contiguous and acc.size == N
    R = olbm._real(dtype)
    lib.orc_d3q7_step(olbm._p(gin), olbm._p(gout), olbm._p(T), *[olbm._p(a) for a in arr[:3]], olbm._p(fl),
                      olbm._p(arr[3]), None if Qa is None else olbm._p(Qa), None if acc is None else olbm._p(acc),
                      C.c_int(nx), C.c_int(ny), C.c_int(nz), *[C.c_int(int(p)) for p in periodic], R(omega),
                      R(omega if omega_s is None else omega_s), R(T_in))
    return gout, T


def face_laplacian(conducting: np.ndarray, k: float):
    """Sparse K with (K T)_c = k sum over conducting face neighbours (T_c - T_nb):
    conduction between conducting cells only (zero flux to every other cell and
    across the domain faces). Returns (K, index of each conducting cell)."""
    import scipy.sparse as sp
    cond = np.asarray(conducting, bool)
    idx = -np.ones(cond.shape, np.int64)
    cells = np.flatnonzero(cond.reshape(-1))
    idx.reshape(-1)[cells] = np.arange(cells.size)
    rows, cols = [], []
    for ax in range(3):
        a = [slice(None)] * 3
        b = [slice(None)] * 3
        a[ax] = slice(0, -1)
        b[ax] = slice(1, None)
        both = cond[tuple(a)] & cond[tuple(b)]
        ia, ib = idx[tuple(a)][both], idx[tuple(b)][both]
        rows += [ia, ib]
        cols += [ib, ia]
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    n = cells.size
    off = sp.coo_matrix((-k * np.ones(r.size), (r, c)), shape=(n, n)).tocsr()
    deg = -np.asarray(off.sum(axis=1)).reshape(-1)
    return (off + sp.diags(deg)).tocsr(), cells


def solid_step(T_old, acc, src, conducting, k_s, rho_cp_s, dt):
    """Implicit conduction step of the solid over ``dt`` flow steps:
    rho_cp_s (T - T_old) + dt K T = acc + dt rho_cp_s src on conducting cells, K the
    face Laplacian with conductance k_s (fluid rho c_p = 1 units; acc the energy the
    fluid handed over in the interval, src the per-step temperature source).
    FP64 sparse direct solve; returns T in T_old's shape (other cells unchanged)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    K, cells = face_laplacian(conducting, k_s)
    shape = np.shape(T_old)
    T_old = np.asarray(T_old, np.float64).reshape(-1)
    rhs = rho_cp_s * T_old[cells] + np.asarray(acc, np.float64).reshape(-1)[cells] \
        + dt * rho_cp_s * np.asarray(src, np.float64).reshape(-1)[cells]
    A = sp.identity(cells.size, format="csr") * rho_cp_s + dt * K
    T = T_old.copy()
    T[cells] = spl.spsolve(A.tocsc(), rhs)
    return T.reshape(shape)


class CoupledLoop:
    """FP64 restatement of the coupled loop: flow step (oracle reference step) ->
    D3Q7 step with that step's velocity -> every ``couple_every`` steps the solid
    conduction step with the accumulated interface energy and the source."""

    def __init__(self, c: dict, plan, conducting, wall, inlet, omega_g, k_s, rho_cp_s, src, couple_every=50,
                 T_in=0.0, dtype=np.float64):
        self.c = c
        self.plan = plan
        self.shape = (c["nx"], c["ny"], c["nz"])
        N = int(np.prod(self.shape))
        self.dtype = dtype
        self.fl = flags(conducting, wall, inlet).reshape(-1)
        self.conducting = np.asarray(conducting, bool)
        self.omega_g, self.k_s, self.rho_cp_s = omega_g, k_s, rho_cp_s
        self.src = np.asarray(src, np.float64).reshape(-1)
        self.every = int(couple_every)
        self.T_in = T_in
        self.f = np.broadcast_to(olbm.W32[:, None].astype(dtype), (19, N)).copy()
        self.T_solid = np.full(N, T_in, np.float64)
        zero = np.zeros(N)
        T0 = np.full(N, T_in)
        self.g = equilibrium(T0, zero, zero, zero, dtype)
        self.acc = np.zeros(N, dtype)
        self.T = T0.astype(dtype)
        self.n = 0

    def run(self, steps: int, on_step: Optional[Callable] = None):
        c = self.c
        for _ in range(steps):
            self.f, rho, u, v, w = olbm.step(self.f, c["solid_flow"], *self.shape, (0, 0, 0), "trt", c["omega"],
                                             c["omega_minus"], None, 0.0, c["links"], None, self.plan)
            self.u = (u, v, w)
            self.g, self.T = step(self.g, self.fl, u, v, w, self.T_solid.astype(self.dtype), self.shape,
                                  omega=self.omega_g, T_in=self.T_in, acc=self.acc)
            self.n += 1
            if self.n % self.every == 0:
                self.T_solid = solid_step(self.T_solid, self.acc, self.src, self.conducting, self.k_s,
                                          self.rho_cp_s, float(self.every))
                self.acc[:] = 0
            if on_step is not None:
                on_step(self)
        return self


def interface_heat(g, fl, shape, solid_kind=CONJ) -> float:
    """Energy per step that crosses from carried-solid nodes into fluid nodes, from
    the post-collision populations g ([7][N]) of a periodic domain: over every
    fluid node x and direction q with x - c_q solid, g*_q(x - c_q) - g*_opp(q)(x)."""
    nx, ny, nz = shape
    g = np.asarray(g).reshape(7, nx, ny, nz)
    kind = (np.asarray(fl).reshape(nx, ny, nz) & 3)
    fluid = kind == FLUID
    solid = kind == solid_kind
    opp = [0, 2, 1, 4, 3, 6, 5]
    e = 0.0
    for q in range(1, 7):
        c = (int(CX7[q]), int(CY7[q]), int(CZ7[q]))
        m = fluid & np.roll(solid, shift=c, axis=(0, 1, 2))  # x fluid, x - c_q solid
        e += float(np.roll(g[q], shift=c, axis=(0, 1, 2))[m].sum() - g[opp[q]][m].sum())
    return e


def interface_links(fl, shape, solid_kind=CONJ) -> int:
    """Fluid-solid faces (lattice links between face neighbours) of a periodic domain."""
    nx, ny, nz = shape
    kind = np.asarray(fl).reshape(nx, ny, nz) & 3
    fluid = kind == FLUID
    solid = kind == solid_kind
    return int(sum((fluid & np.roll(solid, shift=(int(CX7[q]), int(CY7[q]), int(CZ7[q])), axis=(0, 1, 2))).sum()
                   for q in range(1, 7)))


def nusselt(g, T, fl, shape, kappa_f, solid_kind=CONJ) -> dict:
    """Interstitial Nusselt number of a periodic conjugate problem:
    h = Q / (A (mean T_solid - mean T_fluid)), Nu = h d_h / k_f with Q the energy per
    step crossing the interface (interface_heat), A the interface links, d_h =
    4 V_fluid / A and k_f = kappa_f (fluid rho c_p = 1)."""
    kind = np.asarray(fl).reshape(-1) & 3
    T = np.asarray(T, np.float64).reshape(-1)
    Q = interface_heat(g, fl, shape, solid_kind)
    A = interface_links(fl, shape, solid_kind)
    Ts = float(T[kind == solid_kind].mean())
    Tf = float(T[kind == FLUID].mean())
    d_h = 4.0 * float((kind == FLUID).sum()) / A
    h = Q / (A * (Ts - Tf))
    return {"Nu": h * d_h / kappa_f, "h": h, "Q": Q, "A": A, "d_h": d_h, "T_solid_mean": Ts, "T_fluid_mean": Tf}
