"""Pins of the D3Q7 temperature oracle and the solid conduction step (oracle/d3q7.py).

D3Q7 has no counterpart in the reference (its temperature is a finite-volume
operator), so parity is pinned by what the mathematics fixes:

* 1-D advection-diffusion of a Gaussian pulse along each axis, against the
  closed form with the model's diffusivity (1/omega - 1/2)(1/4 - u^2) along the
  flow (the linear equilibrium's known u^2 correction; with u = 0.1 the plain
  (1/omega - 1/2)/4 is measurably wrong, which the test also shows);
* the steady conduction profile between two solids held at different
  temperatures: linear, with the solid temperatures on the faces half-way between
  nodes (anti-bounce-back);
* energy conservation with walls, solids and periodic faces: fluid energy plus the
  energy handed to the solids is constant to rounding;
* the open-boundary energy identity that the full-size coupled run checks on the
  GPU (inlet and outlet plane terms), on an oracle flow;
* the solid step: a Neumann cosine eigenmode of a bar decays by exactly
  1 / (1 + dt lambda / rho_cp), a uniform field without sources is a fixed point,
  and the step's energy balance holds.
"""
import types

import numpy as np
import pytest

from oracle import d3q7 as od
from oracle import lbm as olbm


def _pulse(axis, u_axis, steps=600, n=256, om=1.0 / 0.8, sig0=8.0, z0=80.0):
    shape = [1, 1, 1]
    shape[axis] = n
    N = n
    z = np.arange(n, dtype=np.float64)
    T0 = np.exp(-(z - z0) ** 2 / (2 * sig0 ** 2))
    vel = [np.zeros(N) for _ in range(3)]
    vel[axis][:] = u_axis
    g = od.equilibrium(T0, *vel)
    fl = np.zeros(N, np.uint8)
    for _ in range(steps):
        g, T = od.step(g, fl, *vel, np.zeros(N), tuple(shape), periodic=(1, 1, 1), omega=om)

    def err(kappa):
        s2 = sig0 ** 2 + 2 * kappa * steps
        zc = z0 + u_axis * steps
        ex = sum(sig0 / np.sqrt(s2) * np.exp(-(z - zc - m * n) ** 2 / (2 * s2)) for m in (-1, 0, 1))
        return np.abs(T - ex).max() / ex.max()

    return err, T, T0


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("u", [0.0, 0.1, -0.1])
def test_gaussian_pulse_matches_advection_diffusion(axis, u):
    om = 1.0 / 0.8
    err, T, T0 = _pulse(axis, u, om=om)
    kappa_model = (1 / om - 0.5) * (0.25 - u * u)
    assert err(kappa_model) < 1.5e-3
    if u != 0.0:  # the u^2 term is resolved: the uncorrected diffusivity misses by ~20x more
        assert err((1 / om - 0.5) / 4) > 5e-3
    assert abs(T.sum() - T0.sum()) < 1e-12 * T0.sum()


def test_pulse_transverse_to_the_flow_diffuses_with_the_plain_diffusivity():
    """Diffusion across the flow has no u^2 correction."""
    n, steps, om, sig0, z0 = 128, 400, 1.0 / 0.9, 6.0, 64.0
    z = np.arange(n, dtype=np.float64)
    T0 = np.exp(-(z - z0) ** 2 / (2 * sig0 ** 2))
    N = n
    u = np.full(N, 0.1)
    g = od.equilibrium(T0, u, np.zeros(N), np.zeros(N))
    for _ in range(steps):  # variation along z, flow along x (one node wide, periodic)
        g, T = od.step(g, np.zeros(N, np.uint8), u, np.zeros(N), np.zeros(N), np.zeros(N), (1, 1, n),
                       periodic=(1, 1, 1), omega=om)
    s2 = sig0 ** 2 + 2 * (1 / om - 0.5) / 4 * steps
    ex = sig0 / np.sqrt(s2) * np.exp(-(z - z0) ** 2 / (2 * s2))
    assert np.abs(T - ex).max() / ex.max() < 1.5e-3


def test_linear_profile_between_two_solids():
    ny = 18
    fl = np.zeros((1, ny, 1), np.uint8)
    fl[:, 0] = od.SOLID
    fl[:, -1] = od.SOLID
    Ts = np.zeros(ny)
    Ts[-1] = 1.0
    N = ny
    z = np.zeros(N)
    g = od.equilibrium(np.zeros(N), z, z, z)
    for _ in range(14000):  # slowest mode decays by e every ~390 steps
        g, T = od.step(g, fl, z, z, z, Ts, (1, ny, 1), periodic=(1, 0, 1), omega=1.3)
    j = np.arange(1, ny - 1, dtype=np.float64)
    exact = (j - 0.5) / (ny - 2)  # solid temperatures on the faces at 0.5 and ny - 1.5
    assert np.abs(T[1:-1] - exact).max() < 1e-10
    assert T[0] == 0.0 and T[-1] == 1.0


def test_energy_conserved_with_walls_solids_and_periodic_faces():
    rng = np.random.default_rng(7)
    shape = (12, 10, 14)
    N = int(np.prod(shape))
    kind = rng.choice([0, 0, 0, 0, 1, 2], size=N).astype(np.uint8)
    u, v, w = (0.05 * rng.standard_normal(N) for _ in range(3))
    T0 = rng.random(N)
    Ts = rng.random(N)
    g = od.equilibrium(T0, u, v, w)
    acc = np.zeros(N)
    fluid = kind == 0
    E0 = T0[fluid].sum()
    for s in range(40):
        g, T = od.step(g, kind, u, v, w, Ts, shape, periodic=(1, 1, 1), omega=1.1, acc=acc)
        assert abs(T[fluid].sum() + acc.sum() - E0) < 1e-11 * N
    assert np.abs(acc).max() > 1e-3  # energy did cross the interfaces


def _plan(c):
    from fluxcore3d_amd.solver.lbm_cuda_ext import flow_bc_plan
    return flow_bc_plan(types.SimpleNamespace(cfg=c["cfg"], u_in_lat=c["cfg"].u_in_lat, nx=c["nx"], ny=c["ny"],
                                              nz=c["nz"]))


def open_face_terms(g, fl, shape):
    """Energy the fluid (inlet nodes excluded) gains in the next step through open
    domain faces and loses to inlet nodes, from the post-collision populations g
    ([7][N]); the full-size GPU test uses the same bookkeeping."""
    nx, ny, nz = shape
    g = g.reshape(7, nx, ny, nz)
    fl = fl.reshape(nx, ny, nz)
    F = (fl & 3) == 0
    F_int = F & ((fl & 16) == 0)
    inlet = F & ((fl & 16) != 0)
    e = 0.0
    # open faces: outward population lost, inward population copied in (zero gradient)
    for ax, (q_out_hi, q_out_lo) in enumerate(((1, 2), (3, 4), (5, 6))):
        hi = [slice(None)] * 3
        lo = [slice(None)] * 3
        hi[ax] = -1
        lo[ax] = 0
        m = F_int[tuple(hi)]
        e += (g[q_out_lo][tuple(hi)][m].sum() - g[q_out_hi][tuple(hi)][m].sum())
        m = F_int[tuple(lo)]
        e += (g[q_out_hi][tuple(lo)][m].sum() - g[q_out_lo][tuple(lo)][m].sum())
    # populations sent to inlet nodes are dropped (inlet nodes are reset each step),
    # populations the inlet nodes send are counted through the inlet temperature
    for q, (dx, dy, dz) in enumerate(zip(od.CX7, od.CY7, od.CZ7)):
        if q == 0:
            continue
        tgt = np.roll(inlet, shift=(-dx, -dy, -dz), axis=(0, 1, 2))  # node y with y + c_q an inlet node
        ok = np.ones(shape, bool)  # y + c_q inside the domain
        for ax, d in enumerate((dx, dy, dz)):
            sl = [slice(None)] * 3
            if d == 1:
                sl[ax] = -1
                ok[tuple(sl)] = False
            elif d == -1:
                sl[ax] = 0
                ok[tuple(sl)] = False
        e -= g[q][F_int & tgt & ok].sum()
        src_in = np.roll(inlet, shift=(dx, dy, dz), axis=(0, 1, 2))  # x with x - c_q an inlet node
        ok2 = np.ones(shape, bool)
        for ax, d in enumerate((dx, dy, dz)):
            sl = [slice(None)] * 3
            if d == 1:
                sl[ax] = 0
                ok2[tuple(sl)] = False
            elif d == -1:
                sl[ax] = -1
                ok2[tuple(sl)] = False
        e += np.roll(g[q] * inlet, shift=(dx, dy, dz), axis=(0, 1, 2))[F_int & src_in & ok2].sum()
    return float(e)


def test_open_boundary_energy_identity_on_the_quarter_heat_sink():
    """Fluid energy change = open-face terms - energy handed to the solid, per step,
    with the inlet at a nonzero temperature (so the inlet terms are exercised)."""
    from tests import cases
    c = cases.pin_fin_channel()
    shape = (c["nx"], c["ny"], c["nz"])
    N = int(np.prod(shape))
    cond = c["solid"]
    wall = c["solid_flow"] & ~cond
    inlet = np.zeros(shape, bool)
    inlet[:, :, 0] = True
    fl = od.flags(cond, wall, inlet).reshape(-1)
    F_int = ((fl & 3) == 0) & ((fl & 16) == 0)
    rng = np.random.default_rng(3)
    Ts = np.where(cond.reshape(-1), 1.0 + rng.random(N), 0.0)
    f = np.broadcast_to(olbm.W32[:, None].astype(np.float64), (19, N)).copy()
    T0 = rng.random(N)
    z = np.zeros(N)
    g = od.equilibrium(T0, z, z, z)
    acc = np.zeros(N)
    E = T0[F_int].sum()
    plan = _plan(c)
    for s in range(30):
        f, rho, u, v, w = olbm.step(f, c["solid_flow"], *shape, (0, 0, 0), "trt", c["omega"], c["omega_minus"],
                                    None, 0.0, c["links"], None, plan)
        pred = E + open_face_terms(g, fl, shape) + acc.sum()
        g, T = od.step(g, fl, u, v, w, Ts, shape, omega=1.4, T_in=0.7, acc=acc)
        pred -= acc.sum()
        E = T[F_int].sum()
        assert abs(E - pred) < 1e-11 * N, (s, E, pred)
    assert acc.sum() < 0  # the fluid (random T in [0, 1]) gains from solids at 1..2


def test_solid_step_eigenmode_fixed_point_and_balance():
    n = 24
    k, rcp, dt = 3.0, 1.6, 50.0
    cond = np.zeros((n, 3, 3), bool)
    cond[:, 1, 1] = True  # a bar of n cells along x
    i = np.arange(n)
    mode = np.cos(np.pi * (i + 0.5) / n)  # Neumann eigenvector of the path Laplacian
    lam = 2 * k * (1 - np.cos(np.pi / n))
    T_old = np.zeros(cond.shape)
    T_old[:, 1, 1] = mode
    T = od.solid_step(T_old, np.zeros(cond.shape), np.zeros(cond.shape), cond, k, rcp, dt)
    np.testing.assert_allclose(T[:, 1, 1], mode / (1 + dt * lam / rcp), rtol=1e-12, atol=1e-14)
    # second mode: a wrong conductance or a missing factor changes the ratio between modes
    mode2 = np.cos(2 * np.pi * (i + 0.5) / n)
    lam2 = 2 * k * (1 - np.cos(2 * np.pi / n))
    T_old[:, 1, 1] = mode2
    T = od.solid_step(T_old, np.zeros(cond.shape), np.zeros(cond.shape), cond, k, rcp, dt)
    np.testing.assert_allclose(T[:, 1, 1], mode2 / (1 + dt * lam2 / rcp), rtol=1e-12, atol=1e-14)
    # uniform without sources: fixed point (cells outside the solid untouched)
    T_old = np.where(cond, 2.5, -7.0)
    T = od.solid_step(T_old, np.zeros(cond.shape), np.zeros(cond.shape), cond, k, rcp, dt)
    np.testing.assert_allclose(T, T_old, rtol=1e-13)
    # energy balance on an irregular solid with acc and a source
    rng = np.random.default_rng(1)
    cond = rng.random((9, 8, 7)) < 0.6
    T_old = rng.random(cond.shape)
    acc = rng.standard_normal(cond.shape)
    src = rng.random(cond.shape) * 1e-3
    T = od.solid_step(T_old, acc, src, cond, k, rcp, dt)
    lhs = rcp * (T - T_old)[cond].sum()
    rhs = acc[cond].sum() + dt * rcp * src[cond].sum()
    assert abs(lhs - rhs) < 1e-10 * (abs(rhs) + 1)
    assert np.array_equal(T[~cond], T_old[~cond])


def _fv_periodic_or_walled(kap, src, TL=None, TR=None):
    """Steady 1-D finite volumes with harmonic-mean face conductances 1/(1/(2k_i) +
    1/(2k_j)); walls at -1/2 and n - 1/2 held at TL, TR (conductance 2k), or a
    periodic ring when TL is None (mean fixed at zero)."""
    n = kap.size
    A = np.zeros((n, n))
    b = np.array(src, np.float64)
    pairs = [(i, i + 1) for i in range(n - 1)] + ([(n - 1, 0)] if TL is None else [])
    for i, j in pairs:
        G = 1.0 / (1.0 / (2 * kap[i]) + 1.0 / (2 * kap[j]))
        A[i, i] += G
        A[j, j] += G
        A[i, j] -= G
        A[j, i] -= G
    if TL is not None:
        A[0, 0] += 2 * kap[0]
        b[0] += 2 * kap[0] * TL
        A[-1, -1] += 2 * kap[-1]
        b[-1] += 2 * kap[-1] * TR
        return np.linalg.solve(A, b)
    A = np.vstack([A, np.ones(n)])
    return np.linalg.lstsq(A, np.append(b, 0.0), rcond=None)[0]


def test_conjugate_two_layer_steady_state_is_harmonic_finite_volumes():
    """Fluid (kind 0, u = 0) and carried solid (kind 3, own rate) between two solids
    held at 0 and 1: the steady D3Q7 profile equals finite volumes with harmonic-mean
    face conductances (the face rule of the reference's operator)."""
    n = 30
    fl = np.zeros((n + 2, 1, 1), np.uint8)
    fl[0] = fl[-1] = od.SOLID
    fl[1 + 12:-1] = od.CONJ
    Ts = np.zeros(n + 2)
    Ts[-1] = 1.0
    kf, ks = 0.05, 0.4
    om_f, om_s = 1 / (0.5 + 4 * kf), 1 / (0.5 + 4 * ks)
    N = n + 2
    z = np.zeros(N)
    g = od.equilibrium(np.zeros(N), z, z, z)
    for _ in range(60000):
        g, T = od.step(g, fl, z, z, z, Ts, (N, 1, 1), periodic=(0, 1, 1), omega=om_f, omega_s=om_s)
    kap = np.where(fl[1:-1, 0, 0] == od.CONJ, ks, kf)
    exact = _fv_periodic_or_walled(kap, np.zeros(n), 0.0, 1.0)
    assert np.abs(T[1:-1] - exact).max() < 1e-11


def test_conjugate_periodic_tiling_is_exact():
    rng = np.random.default_rng(4)
    base = (6, 5, 7)
    Nb = int(np.prod(base))
    kind = rng.choice([0, 0, 0, 2, 3, 3], size=base).astype(np.uint8)
    fields = [0.05 * rng.standard_normal(base) for _ in range(3)] + [rng.random(base), 1e-3 * rng.standard_normal(base)]
    tiled = [np.tile(a, (2, 2, 2)) for a in [kind] + fields]
    out = []
    for kd, (u, v, w, T0, Q) in ((kind, fields), (tiled[0], tiled[1:])):
        shape = kd.shape
        g = od.equilibrium(T0, u, v, w)
        for _ in range(25):
            g, T = od.step(g, kd, u, v, w, np.zeros(kd.size), shape, periodic=(1, 1, 1), omega=1.3, omega_s=0.7,
                           Q=Q)
        out.append(T.reshape(shape))
    assert np.array_equal(np.tile(out[0], (2, 2, 2)), out[1])
    assert Nb > 0


def test_energy_with_carried_solid_and_sources():
    rng = np.random.default_rng(9)
    shape = (10, 9, 8)
    N = int(np.prod(shape))
    kind = rng.choice([0, 0, 2, 3], size=N).astype(np.uint8)
    u, v, w = (0.04 * rng.standard_normal(N) for _ in range(3))
    T0 = rng.random(N)
    Q = 1e-3 * rng.standard_normal(N)
    live = (kind == 0) | (kind == 3)
    g = od.equilibrium(T0, u, v, w)
    for s in range(1, 31):
        g, T = od.step(g, kind, u, v, w, np.zeros(N), shape, periodic=(1, 1, 1), omega=1.2, omega_s=0.8, Q=Q)
        # the source enters the populations after the temperature is read
        assert abs(T[live].sum() - (T0[live].sum() + (s - 1) * Q[live].sum())) < 1e-11 * N


def test_interface_heat_of_the_periodic_slab_equals_the_source():
    """Periodic slab (carried solid a nodes, fluid b nodes), source q in the solid and
    the balancing sink in the fluid: at steady state the energy crossing the two
    interfaces per step, read from the populations, is q a."""
    a, b = 7, 11
    n = a + b
    fl = np.zeros((n, 1, 1), np.uint8)
    fl[:a] = od.CONJ
    kf, ks, q = 0.05, 0.3, 1e-3
    src = np.where(fl.reshape(-1) == od.CONJ, q, -q * a / b)
    z = np.zeros(n)
    g = od.equilibrium(np.zeros(n), z, z, z)
    for _ in range(40000):
        g, T = od.step(g, fl, z, z, z, z, (n, 1, 1), periodic=(1, 1, 1), omega=1 / (0.5 + 4 * kf),
                       omega_s=1 / (0.5 + 4 * ks), Q=src)
    assert abs(od.interface_heat(g, fl, (n, 1, 1)) - q * a) < 1e-12
    assert od.interface_links(fl, (n, 1, 1)) == 2


def test_nusselt_of_a_three_layer_slab_matches_finite_volumes():
    """Solid held at 1 | carried solid (a nodes) | fluid (b nodes) | solid held at 0,
    no sources: the steady state is harmonic-mean finite volumes, so the heat
    through the carried-solid/fluid face, both layer means and the Nusselt number
    J / (A (T_s - T_f)) d_h / k_f follow from the finite-volume solution."""
    a, b = 9, 13
    n = a + b
    fl = np.zeros((n + 2, 1, 1), np.uint8)
    fl[0] = fl[-1] = od.SOLID
    fl[1:1 + a] = od.CONJ
    Ts = np.zeros(n + 2)
    Ts[0] = 1.0
    kf, ks = 0.05, 0.35
    N = n + 2
    z = np.zeros(N)
    g = od.equilibrium(np.zeros(N), z, z, z)
    for _ in range(60000):
        g, T = od.step(g, fl, z, z, z, Ts, (N, 1, 1), periodic=(0, 1, 1), omega=1 / (0.5 + 4 * kf),
                       omega_s=1 / (0.5 + 4 * ks))
    res = od.nusselt(g, T, fl, (N, 1, 1), kf)
    kap = np.where(fl[1:-1, 0, 0] == od.CONJ, ks, kf)
    Tfv = _fv_periodic_or_walled(kap, np.zeros(n), 1.0, 0.0)
    J = (1.0 - Tfv[0]) * 2 * ks  # wall face of the first node
    assert abs(res["Q"] - J) < 1e-11 and res["A"] == 1 and res["d_h"] == 4.0 * b
    assert abs(res["T_solid_mean"] - Tfv[:a].mean()) < 1e-11 and abs(res["T_fluid_mean"] - Tfv[a:].mean()) < 1e-11
    Nu = J / (Tfv[:a].mean() - Tfv[a:].mean()) * 4.0 * b / kf
    assert abs(res["Nu"] - Nu) < 1e-9 * Nu


def test_gyroid_flow_and_heat_on_96_cube_are_the_tiled_period():
    """The 96^3 crop of config 4 is an exact sub-problem: the oracle's MRT flow step
    (Bouzidi links, body force) and the conjugate D3Q7 step on the 96^3 periodic
    gyroid equal the 48^3 period tiled, bit for bit."""
    from tests import cases
    p = cases.gyroid_heat_params()
    out = {}
    for n in (48, 96):
        c = cases.gyroid_box(n=n, period=48.0)
        N = n ** 3
        solid = c["solid"].reshape(-1)
        fl = np.where(solid, od.CONJ, od.FLUID).astype(np.uint8)
        n_s = int(solid.sum())
        Q = np.where(solid, p["q_solid"], -p["q_solid"] * n_s / (N - n_s))
        f = np.broadcast_to(olbm.W32[:, None].astype(np.float64), (19, N)).copy()
        z = np.zeros(N)
        g = od.equilibrium(z, z, z, z)
        kf = p["kappa_f"]
        for _ in range(3):
            f, rho, u, v, w = olbm.step(f, c["solid"], n, n, n, (1, 1, 1), "mrt", p["omega"], 1.0,
                                        p["S"].astype(np.float64), 0.0, c["links"], force=p["force"])
            g, T = od.step(g, fl, u, v, w, z, (n, n, n), periodic=(1, 1, 1), omega=1 / (0.5 + 4 * kf),
                           omega_s=1 / (0.5 + 4 * p["k_ratio"] * kf), Q=Q)
        out[n] = [a.reshape(n, n, n) for a in (u, v, w, T)] + [f.reshape(19, n, n, n)]
    for a, b in zip(out[48][:4], out[96][:4]):
        assert np.array_equal(np.tile(a, (2, 2, 2)), b)
    assert np.array_equal(np.tile(out[48][4], (1, 2, 2, 2)), out[96][4])


@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 0, 1), (1, 1, 1)])
def test_neighbour_mask_against_brute_force(periodic):
    """The per-node mask the D3Q7 step reads against a loop over every node and direction:
    on solid nodes solved apart, bit q - 1 when x + c_q is inside the domain and a fluid
    node off the inlet; elsewhere, when x - c_q is inside and a wall or such a solid."""
    import torch
    from fluxcore3d_amd.solver.cht_d3q7 import neighbour_mask, FLUID, SOLID, WALL, CONJ, INLET

    rng = np.random.default_rng(7)
    shape = (5, 6, 7)
    fl = rng.choice(np.array([FLUID, SOLID, WALL, CONJ, INLET], np.uint8), size=shape)
    got = neighbour_mask(torch.from_numpy(fl), periodic).numpy()
    c = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    want = np.zeros(shape, np.uint8)
    for i, j, k in np.ndindex(*shape):
        solid = (fl[i, j, k] & 3) == SOLID
        for q, d in enumerate(c):
            sgn = 1 if solid else -1  # a solid looks at x + c_q, every other node at x - c_q
            p = [i + sgn * d[0], j + sgn * d[1], k + sgn * d[2]]
            inside = True
            for a in range(3):
                if not 0 <= p[a] < shape[a]:
                    if periodic[a]:
                        p[a] %= shape[a]
                    else:
                        inside = False
            if not inside:
                continue
            nb = fl[tuple(p)]
            if solid and (nb & 3) == FLUID and not (nb & INLET):
                want[i, j, k] |= 1 << q
            if not solid and (nb & 3) in (SOLID, WALL):
                want[i, j, k] |= 1 << q
    assert got.dtype == np.uint8
    np.testing.assert_array_equal(got, want)
