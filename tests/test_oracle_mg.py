"""Pins of the multigrid oracle (oracle/mg.py) — CPU only."""
import numpy as np
import pytest

from oracle import mg
from tests import thermal_cases


def _rand_fields(shape, seed=0, dtype=np.float64, flow=True):
    rng = np.random.default_rng(seed)
    k = (0.1 + rng.random(shape)).astype(dtype)
    rcp = (0.5 + rng.random(shape)).astype(dtype)
    u, v, w = ((0.05 * rng.standard_normal(shape)) if flow else np.zeros(shape) for _ in range(3))
    return k, rcp, u.astype(dtype), v.astype(dtype), w.astype(dtype)


def _brute_operator(k, rcp, u, v, w, ax, out_idx, dx, s, dt):
    """Dense A = I + dt L assembled face by face from the definition
    (harmonic face conductivity, first-order upwind, ghost rules)."""
    shape = k.shape
    n = k.size
    A = np.eye(n)
    idx = np.arange(n).reshape(shape)
    for c in np.ndindex(shape):
        row = idx[c]
        for dim in range(3):
            for step in (+1, -1):
                nb = list(c)
                if step > 0:
                    nb[dim] = c[dim] + 1 if c[dim] < shape[dim] - 1 else c[dim] - 1
                else:
                    if c[dim] > 0:
                        nb[dim] = c[dim] - 1
                    else:
                        nb[dim] = 0 if (ax == dim and out_idx != 0) else 1
                nb = tuple(nb)
                kh = 2 * k[c] * k[nb] / (k[c] + k[nb] + 1e-30)
                vel = (u, v, w)[dim][c]
                up = max(-vel, 0.0) if step > 0 else max(vel, 0.0)
                a = s / dx**2 * kh / (rcp[c] + 1e-30) + s / dx * up
                A[row, row] += dt * a
                A[row, idx[nb]] -= dt * a
    return A


@pytest.mark.parametrize("ax,sign", [(0, 1), (1, -1), (2, 1)])
def test_operator_against_brute_force(ax, sign):
    shape = (4, 3, 5)
    k, rcp, u, v, w = _rand_fields(shape, seed=ax)
    out_idx = shape[ax] - 1 if sign == 1 else 0
    co = mg.build_coeffs_full(k, rcp, u, v, w, ax, out_idx, 2.0, 1.5)
    A = mg.dense_operator(co, 0.3, ax, out_idx)
    B = _brute_operator(k, rcp, u, v, w, ax, out_idx, 2.0, 1.5, 0.3)
    assert np.abs(A - B).max() < 1e-12


def test_prolong_matches_torch_interpolate():
    torch = pytest.importorskip("torch")
    import torch.nn.functional as F
    rng = np.random.default_rng(2)
    for cs, fs in (((4, 3, 5), (8, 6, 10)), ((4, 3, 5), (9, 7, 11)), ((8, 4, 4), (16, 8, 8))):
        xc = rng.standard_normal(cs).astype(np.float32)
        got = mg.prolong_2x(xc, *fs)
        want = F.interpolate(torch.from_numpy(xc)[None, None], size=fs, mode="trilinear",
                             align_corners=False)[0, 0].numpy()
        assert np.abs(got - want).max() < 2e-6


def test_restrict_is_block_mean():
    x = np.random.default_rng(3).standard_normal((5, 4, 7))
    r = mg.restrict_2x(x)
    assert r.shape == (2, 2, 3)
    for c in np.ndindex(r.shape):
        blk = x[2 * c[0]:2 * c[0] + 2, 2 * c[1]:2 * c[1] + 2, 2 * c[2]:2 * c[2] + 2]
        assert abs(r[c] - blk.mean()) < 1e-14


def _spsolve_masked(h, b):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    lv = h.levels[0]
    A = mg.dense_operator(lv["co"], h.dt, h.ax, lv["out_idx"])
    m = lv["mask"].ravel().astype(bool)
    x = np.zeros(b.size)
    x[m] = spla.spsolve(sp.csr_matrix(A[np.ix_(m, m)]), (b.ravel() * m)[m])
    return x.reshape(b.shape)


def test_vcycle_solve_and_fcg_reach_direct_solution():
    c = thermal_cases.sphere_conduction(16, 8, 8, dt=50.0)
    h = mg.Hierarchy(c["k"].astype(np.float64), c["rcp"].astype(np.float64), c["u"], c["v"], c["w"], c["mask"], 0, 1,
                     1.0, omega=0.6)
    h.set_dt(c["dt"])
    b = c["b"].astype(np.float64)
    x_direct = _spsolve_masked(h, b)
    # solve() measures ||b - Lo|| without the mask (nlevel_mg_cuda.py:133-135),
    # so Dirichlet rows leave a floor; the masked residual and the solution converge
    x, hist = h.solve(b, max_cycles=300, tol=1e-10)
    rm = (b - h.mop(x, 0)) * h.levels[0]["mask"]
    assert np.linalg.norm(rm) / np.linalg.norm(b) < 1e-9
    assert np.abs(x - x_direct).max() / np.abs(x_direct).max() < 1e-8
    xf, info = mg.fcg_mg_fp64(h, b, tol=1e-10, max_it=100)
    assert info["converged"] and info["it"] < len(hist)
    assert np.abs(xf - x_direct).max() / np.abs(x_direct).max() < 1e-8
    xp, pinfo = mg.pcg_mg(h, b, tol=1e-8, max_it=100)
    assert pinfo["res_rel"] < 1e-6


def test_hierarchy_levels_and_masks():
    c = thermal_cases.sphere_conduction(64, 32, 32)
    h = mg.Hierarchy(c["k"], c["rcp"], c["u"], c["v"], c["w"], c["mask"], 0, 1, 1.0)
    assert [lv["shape"] for lv in h.levels] == [(64, 32, 32), (32, 16, 16), (16, 8, 8)]
    # a coarse cell is an unknown only when all 8 children are
    m1 = h.levels[1]["mask"]
    m0 = c["mask"].astype(np.float32)
    assert np.array_equal(m1 > 0, mg.restrict_2x(m0) >= 0.999)


def test_bicgstab_reaches_direct_solution_on_advection_system():
    """The oracle's BiCGSTAB (mg_pcg.py:103-175) on a non-symmetric system (upwind
    advection, varying rho*c_p, walls held at 0, two levels), FP64 throughout, reaches
    the spsolve solution; the recursive residual it reports equals the true masked
    residual of the returned x. In FP32, as the reference runs it, it reaches 1e-5."""
    rng = np.random.default_rng(4)
    shape = (12, 10, 14)
    k = 0.1 + rng.random(shape)
    rcp = 0.5 + rng.random(shape)
    u, v, w = (0.3 * rng.standard_normal(shape) for _ in range(3))
    mask = np.ones(shape, bool)
    mask[:, 0, :] = mask[:, -1, :] = False
    b = rng.standard_normal(shape) * mask
    h = mg.Hierarchy(k, rcp, u, v, w, mask, 2, 1, 1.5, omega=0.6)
    h.set_dt(0.3)
    assert len(h.levels) == 2
    x_direct = _spsolve_masked(h, b)
    x, info = mg.bicgstab_mg(h, b, tol=1e-11, max_it=200)
    assert info["converged"], info["hist"][-3:]
    assert np.abs(x - x_direct).max() / np.abs(x_direct).max() < 1e-9
    rt = (b - h.mop(x, 0)) * h.levels[0]["mask"]
    x2, info2 = mg.bicgstab_mg(h, b, x0=x, tol=1e-30, max_it=1)
    assert abs(info2["hist"][0] - np.linalg.norm(rt) / np.linalg.norm(b)) < 1e-12
    h32 = mg.Hierarchy(*(a.astype(np.float32) for a in (k, rcp, u, v, w)), mask, 2, 1, 1.5, omega=0.6)
    h32.set_dt(0.3)
    x32, info32 = mg.bicgstab_mg(h32, b.astype(np.float32), tol=1e-5, max_it=200)
    assert info32["converged"] and np.abs(x32 - x_direct).max() / np.abs(x_direct).max() < 1e-3


@pytest.mark.parametrize("ax,out_idx", [(0, 8), (0, 0), (2, 4)])
def test_reconstruct_matches_face_loop(ax, out_idx):
    """The vectorised reconstruction against a per-cell loop over the faces in
    priority x, y, z (thermal_kernels.cu:61-114)."""
    rng = np.random.default_rng(1)
    n = (9, 7, 5)
    T0, th = (rng.standard_normal(n).astype(np.float32) for _ in range(2))
    val = T0 + th
    want = val.copy()
    for c in np.ndindex(*n):
        src = list(c)
        for a in range(3):
            if c[a] == n[a] - 1:
                src[a] = c[a] - 1
                break
            if c[a] == 0 and (a != ax or out_idx == 0):
                src[a] = 1
                break
        want[c] = val[tuple(src)]
    np.testing.assert_array_equal(mg.reconstruct(T0, th, ax, out_idx), want)


def _uniform_box(nx=6, ny=9, nz=5, kval=2.0):
    k = np.full((nx, ny, nz), kval)
    z = np.zeros_like(k)
    return mg.Hierarchy(k, np.ones_like(k), z, z, z, np.ones_like(k, dtype=bool), 0, 1, 1.0, omega=0.8,
                        min_coarse_cells=1)


@pytest.mark.parametrize("n_steps", [1, 2, 3])
def test_chebyshev_smooth_fixed_point_and_eigenmode(n_steps):
    """chebyshev_smooth (mg_pcg.py:195-247), FP64: the exact solution of A x = b is a
    fixed point, and on an eigenvector v of A (uniform conduction box, v = cos(pi a j /
    (ny - 1)) across the transverse y axis, where the mirror ghosts make it exact) the
    error is multiplied by the smoother's polynomial in the eigenvalue m of invD A,
    m = omega (1 + 2 a_y dt (1 - cos theta)) / (1 + dt diagL)."""
    h = _uniform_box()
    h.set_dt(0.7)
    nx, ny, nz = h.levels[0]["shape"]
    rng = np.random.default_rng(1)
    x_star = rng.standard_normal((nx, ny, nz))
    b = h.mop(x_star, 0)
    np.testing.assert_allclose(mg.chebyshev_smooth(h, 0, x_star, b, n_steps), x_star, atol=1e-12)
    co = h.levels[0]["co"]
    a_y, diag = float(co[2][0, 2, 0]), float(co[6][0, 2, 0])
    theta = np.pi * 3 / (ny - 1)
    v = np.broadcast_to(np.cos(theta * np.arange(ny))[None, :, None], (nx, ny, nz)).copy()
    np.testing.assert_allclose(h.mop(v, 0), (1 + 0.7 * 2 * a_y * (1 - np.cos(theta))) * v, atol=1e-12)
    m = 0.8 * (1 + 0.7 * 2 * a_y * (1 - np.cos(theta))) / (1 + 0.7 * diag)
    lam_max = 1 + 0.7 * diag
    d, c = 0.55 * lam_max, 0.45 * lam_max
    e, p, rho_prev = 1 - m, -m, 1.0  # scalar error recurrence of the smoother
    for _ in range(1, n_steps):
        rho = 1.0 / (2 * d / c - rho_prev)
        p = rho * (2 * d / c * (-m * e) + rho_prev * p)
        e, rho_prev = e + p, rho
    got = mg.chebyshev_smooth(h, 0, x_star + v, b, n_steps) - x_star
    np.testing.assert_allclose(got, e * v, atol=1e-12)
