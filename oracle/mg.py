"""NumPy restatement of the reference's geometric multigrid thermal solver.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py). Follows
solver/nlevel_mg_cuda.py, solver/mg_kernels.cu and solver/mg_pcg.py; every
operation keeps the dtype of its inputs (float32 for parity, float64 for the
FP64 oracle) and the operation order of the reference.
"""
from __future__ import annotations

import numpy as np


def _shift(A, dim, step, ax, out_idx):
    """Neighbour field with the reference's boundary ghosts (nlevel_mg_cuda.py:431-471):
    high face mirrors to n-2; low face mirrors to 1 except on the flow axis with
    the inlet at index 0, where it copies itself."""
    B = A.copy()
    sl = lambda a, b: tuple(slice(a, b) if d == dim else slice(None) for d in range(3))  # noqa: E731
    idx = lambda i: tuple(i if d == dim else slice(None) for d in range(3))  # noqa: E731
    if step == +1:
        B[sl(None, -1)] = A[sl(1, None)]
        B[idx(-1)] = A[idx(-2)]
    else:
        B[sl(1, None)] = A[sl(None, -1)]
        if (ax == dim and out_idx == 0) or ax != dim:
            B[idx(0)] = A[idx(1)]
        else:
            B[idx(0)] = A[idx(0)]
    return B


def build_coeffs_full(k, rho_cp, u, v, w, ax, out_idx, dx, s):
    """6 face coefficients + diagL (nlevel_mg_cuda.py:418-507): harmonic-mean
    diffusion s/dx^2 * harm(k)/rho_cp plus first-order upwind s/dx * u."""
    dt = k.dtype.type
    eps = dt(1e-30)
    inv_rcp = dt(1.0) / (rho_cp + eps)
    kE, kW = _shift(k, 0, +1, ax, out_idx), _shift(k, 0, -1, ax, out_idx)
    kN, kS = _shift(k, 1, +1, ax, out_idx), _shift(k, 1, -1, ax, out_idx)
    kU, kD = _shift(k, 2, +1, ax, out_idx), _shift(k, 2, -1, ax, out_idx)
    harm = lambda a, b: dt(2.0) * a * b / (a + b + eps)  # noqa: E731
    cd = dt(s / (dx * dx))
    ca = dt(s / dx)
    aE = cd * harm(k, kE) * inv_rcp + ca * np.maximum(-u, 0)
    aW = cd * harm(k, kW) * inv_rcp + ca * np.maximum(u, 0)
    aN = cd * harm(k, kN) * inv_rcp + ca * np.maximum(-v, 0)
    aS = cd * harm(k, kS) * inv_rcp + ca * np.maximum(v, 0)
    aU = cd * harm(k, kU) * inv_rcp + ca * np.maximum(-w, 0)
    aD = cd * harm(k, kD) * inv_rcp + ca * np.maximum(w, 0)
    diagL = aE + aW + aN + aS + aU + aD
    return aE, aW, aN, aS, aU, aD, diagL


def neighbours(T, ax, out_idx):
    """_fill_nb_into (nlevel_mg_cuda.py:376-413) == mop_fused ghosts (mg_kernels.cu:37-44,74-90)."""
    return (_shift(T, 0, +1, ax, out_idx), _shift(T, 0, -1, ax, out_idx), _shift(T, 1, +1, ax, out_idx),
            _shift(T, 1, -1, ax, out_idx), _shift(T, 2, +1, ax, out_idx), _shift(T, 2, -1, ax, out_idx))


def L_apply(T, co, ax, out_idx):
    """L(T) = diagL T - sum a_nb T_nb in the kernel's order (mg_kernels.cu:99-103)."""
    aE, aW, aN, aS, aU, aD, diagL = co
    TE, TW, TN, TS, TU, TD = neighbours(T, ax, out_idx)
    return diagL * T - aE * TE - aW * TW - aN * TN - aS * TS - aU * TU - aD * TD


def mop(T, co, dt, ax, out_idx):
    """Lo = T + dt L(T) (mg_kernels.cu:55-106)."""
    return T + T.dtype.type(dt) * L_apply(T, co, ax, out_idx)


def restrict_2x(x):
    """2x2x2 mean with odd extents truncated (nlevel_mg_cuda.py:512-528)."""
    n0, n1, n2 = (d // 2 * 2 for d in x.shape)
    v = x[:n0, :n1, :n2].reshape(n0 // 2, 2, n1 // 2, 2, n2 // 2, 2)
    return v.mean(axis=(1, 3, 5), dtype=x.dtype).astype(x.dtype)


def _lin_index(n_out, n_in, dtype):
    """Source index and weight of linear interpolation, align_corners=False."""
    scale = dtype(n_in) / dtype(n_out)
    src = scale * (np.arange(n_out, dtype=dtype) + dtype(0.5)) - dtype(0.5)
    src = np.maximum(src, dtype(0))
    i0 = np.floor(src).astype(np.int64)
    i0 = np.minimum(i0, n_in - 1)
    i1 = np.minimum(i0 + 1, n_in - 1)
    l1 = (src - i0.astype(dtype)).astype(dtype)
    return i0, i1, (dtype(1) - l1).astype(dtype), l1


def prolong_2x(xc, nx, ny, nz):
    """Trilinear interpolation to (nx, ny, nz), align_corners=False
    (F.interpolate as used by nlevel_mg_cuda.py:531-537)."""
    dt = xc.dtype.type
    out = xc
    for axis, n in enumerate((nx, ny, nz)):
        i0, i1, l0, l1 = _lin_index(n, out.shape[axis], dt)
        shp = [1, 1, 1]
        shp[axis] = n
        out = np.take(out, i0, axis=axis) * l0.reshape(shp) + np.take(out, i1, axis=axis) * l1.reshape(shp)
    return out.astype(xc.dtype)


class Hierarchy:
    """Levels of NLevelGeometricMGSolver._build_hierarchy (nlevel_mg_cuda.py:188-260)."""

    def __init__(self, k, rho_cp, u, v, w, mask, ax, sign, s, omega=0.8, n_pre=2, n_post=2, n_coarse=20,
                 min_coarse_cells=4):
        self.ax, self.sign, self.s = ax, sign, s
        self.omega, self.n_pre, self.n_post, self.n_coarse = omega, n_pre, n_post, n_coarse
        dt = k.dtype
        self.levels = []
        kl, rl, ul, vl, wl, ml = k, rho_cp, u, v, w, mask
        while True:
            nx, ny, nz = kl.shape
            n_ax = (nx, ny, nz)[ax]
            out_idx = n_ax - 1 if sign == +1 else 0
            dx = float(2 ** len(self.levels))
            co = build_coeffs_full(kl, rl, ul, vl, wl, ax, out_idx, dx, s)
            self.levels.append({"shape": (nx, ny, nz), "out_idx": out_idx, "co": co, "mask": ml.astype(dt)})
            if min(nx // 2, ny // 2, nz // 2) <= min_coarse_cells:
                break
            kl, rl, ul, vl, wl = (restrict_2x(a) for a in (kl, rl, ul, vl, wl))
            ml = restrict_2x(ml.astype(dt)) >= 0.999
        self.dt = None

    def set_dt(self, dt):
        """_update_inv_diag (nlevel_mg_cuda.py:262-267)."""
        self.dt = dt
        for lv in self.levels:
            d = lv["co"][6].dtype.type
            diagM = np.maximum(d(1.0) + d(dt) * lv["co"][6], d(1e-12))
            lv["invD"] = (d(1.0) / diagM) * d(self.omega) * lv["mask"]

    def mop(self, T, level):
        lv = self.levels[level]
        return mop(T, lv["co"], self.dt, self.ax, lv["out_idx"])

    def smooth(self, x, b, level):
        lv = self.levels[level]
        return x + lv["invD"] * (b - self.mop(x, level))

    def vcycle(self, level, x, b):
        """_vcycle (nlevel_mg_cuda.py:271-307)."""
        lv = self.levels[level]
        if level == len(self.levels) - 1:
            for _ in range(self.n_coarse):
                x = self.smooth(x, b, level)
            return x
        for _ in range(self.n_pre):
            x = self.smooth(x, b, level)
        r = (b - self.mop(x, level)) * lv["mask"]
        bc = restrict_2x(r)
        xc = self.vcycle(level + 1, np.zeros_like(bc), bc)
        e = prolong_2x(xc, *lv["shape"])
        x = (x + x.dtype.type(0.75) * e) * lv["mask"]
        for _ in range(self.n_post):
            x = self.smooth(x, b, level)
        return x

    def precond(self, r):
        """_precond (mg_pcg.py:178-187): one V-cycle from zero."""
        m = self.levels[0]["mask"]
        z = self.vcycle(0, np.zeros_like(r), r * m)
        return z * m

    def solve(self, b, x0=None, max_cycles=50, tol=1e-6):
        """Stand-alone V-cycle iteration (nlevel_mg_cuda.py:102-162)."""
        m = self.levels[0]["mask"]
        x = np.zeros_like(b) if x0 is None else x0 * m
        bb = b * m
        bn = float(np.linalg.norm(b)) + 1e-30
        hist = []
        for cyc in range(max_cycles):
            x = self.vcycle(0, x, bb)
            r = bb - self.mop(x, 0)
            res = float(np.linalg.norm(r)) / bn
            hist.append(res)
            if not res < float("inf") or (cyc > 2 and res > 10.0 * hist[0]) or res < tol:
                break
        return x * m, hist


def pcg_mg(h: Hierarchy, b, x0=None, tol=1e-3, max_it=15):
    """pcg_mg (mg_pcg.py:33-97): CG preconditioned by one V-cycle; tolerance on
    the recursively updated ||r|| / ||b||."""
    m = h.levels[0]["mask"]
    x = np.zeros_like(b) if x0 is None else x0 * m
    r = (b - h.mop(x, 0)) * m
    z = h.precond(r)
    p = z.copy()
    rz = float(np.dot(r.ravel(), z.ravel()))
    bnorm = float(np.dot(b.ravel(), b.ravel())) ** 0.5 + 1e-30
    converged, it, hist = False, 0, []
    for it in range(max_it):
        res = float(np.dot(r.ravel(), r.ravel())) ** 0.5 / bnorm
        hist.append(res)
        if res < tol:
            converged = True
            break
        Ap = h.mop(p, 0)
        pAp = float(np.dot(p.ravel(), Ap.ravel()))
        if abs(pAp) < 1e-30:
            break
        alpha = rz / pAp
        x = x + x.dtype.type(alpha) * p
        r = (r - r.dtype.type(alpha) * Ap) * m
        z = h.precond(r)
        rz_new = float(np.dot(r.ravel(), z.ravel()))
        beta = rz_new / (rz + 1e-30)
        p = z + p.dtype.type(beta) * p
        rz = rz_new
    x = x * m
    return x, {"converged": converged, "it": it + 1, "res_rel": hist[-1] if hist else None, "hist": hist}


def bicgstab_mg(h: Hierarchy, b, x0=None, tol=1e-3, max_it=15, outer=None):
    """bicgstab_mg (mg_pcg.py:103-175): BiCGSTAB preconditioned by one V-cycle, with the
    reference's masking (v = A p_hat and t = A s_hat keep the masked rows; s, r and x
    are masked) and its 1e-30 guards. ``outer`` is the dtype of the vectors and dot
    products (default: the hierarchy's, as in the reference); the V-cycle sees its
    input cast to the hierarchy dtype."""
    hd = h.levels[0]["co"][0].dtype
    od = np.dtype(outer) if outer is not None else hd
    t_ = od.type
    m = h.levels[0]["mask"].astype(od)
    co = tuple(c.astype(od) for c in h.levels[0]["co"])
    A = lambda v: mop(v, co, h.dt, h.ax, h.levels[0]["out_idx"])  # noqa: E731  (unmasked Lo, as _Mop)
    P = lambda v: h.precond(v.astype(hd)).astype(od)  # noqa: E731
    dot = lambda a, c: float(np.dot(a.ravel(), c.ravel()))  # noqa: E731
    b = b.astype(od)
    x = np.zeros_like(b) if x0 is None else x0.astype(od) * m
    r = (b - A(x)) * m
    r_hat = r.copy()
    bnorm = dot(b, b) ** 0.5 + 1e-30
    rho = alpha = omega = 1.0
    v = np.zeros_like(b)
    p = np.zeros_like(b)
    converged, it, hist = False, 0, []
    for it in range(max_it):
        res = dot(r, r) ** 0.5 / bnorm
        hist.append(res)
        if res < tol:
            converged = True
            break
        rho_new = dot(r_hat, r)
        if abs(rho_new) < 1e-30:
            break
        beta = (rho_new / (rho + 1e-30)) * (alpha / (omega + 1e-30))
        p = (t_(beta) * p + r) - t_(beta * omega) * v
        p_hat = P(p)
        v = A(p_hat)
        alpha = rho_new / (dot(r_hat, v) + 1e-30)
        s = (r - t_(alpha) * v) * m
        if dot(s, s) ** 0.5 / bnorm < tol:
            x = x + t_(alpha) * p_hat
            converged = True
            break
        s_hat = P(s)
        t = A(s_hat)
        omega = dot(t, s) / (dot(t, t) + 1e-30)
        x = ((x + t_(alpha) * p_hat) + t_(omega) * s_hat) * m
        r = (s - t_(omega) * t) * m
        rho = rho_new
    x = x * m
    return x, {"converged": converged, "it": it + 1, "res_rel": dot(r, r) ** 0.5 / bnorm, "hist": hist}


def reconstruct(T0, theta, ax, out_idx):
    """thermal_reconstruct (thermal_kernels.cu:61-114): T = T0 + theta, then every face
    cell copies its inner neighbour (Neumann), faces taken in the priority x, y, z; the
    low face of the flow axis is the inlet (Dirichlet) unless out_idx == 0."""
    val = T0 + theta
    out = val.copy()
    n = val.shape
    done = np.zeros(n, bool)
    for a in range(3):
        hi = [slice(None)] * 3
        hi[a] = n[a] - 1
        src_hi = [slice(None)] * 3
        src_hi[a] = n[a] - 2
        sel = ~done[tuple(hi)]
        out[tuple(hi)][sel] = val[tuple(src_hi)][sel]
        done[tuple(hi)] = True
        if a != ax or out_idx == 0:
            lo = [slice(None)] * 3
            lo[a] = 0
            src_lo = [slice(None)] * 3
            src_lo[a] = 1
            sel = ~done[tuple(lo)]
            out[tuple(lo)][sel] = val[tuple(src_lo)][sel]
            done[tuple(lo)] = True
    return out


def thermal_step(h: Hierarchy, T_old, T0, src, mask, dt, tol=1e-3, max_cycles=20):
    """One implicit pseudo-time step of the fused thermal driver (thermal_cuda_ext.py:
    55-201): LT0 = L(T0) (L_apply_fine), rhs = mask ((T_old + dt src) - (T0 + dt LT0)),
    theta0 = mask (T_old - T0), theta from the V-cycle iteration (solve), theta *= mask,
    T = reconstruct(T0 + theta). ``h`` must have set_dt(dt)."""
    d = T_old.dtype.type
    lv = h.levels[0]
    LT0 = (h.mop(T0, 0) - T0) / d(h.dt)
    rhs = mask * ((T_old + d(dt) * src) - (T0 + d(dt) * LT0))
    theta0 = mask * (T_old - T0)
    theta, hist = h.solve(rhs, x0=theta0, max_cycles=max_cycles, tol=tol)
    theta = theta * mask
    return reconstruct(T0, theta, h.ax, lv["out_idx"]), hist


def thermal_step_guarded(h: Hierarchy, T_old, T0, src, mask, dt, T_lo, T_hi, tol=1e-3, max_cycles=20):
    """thermal_step with the fused driver's validity check (thermal_cuda_ext.py:146-160,
    physics_check_kernel thermal_kernels.cu:175-200): the candidate T0 + theta (theta
    masked) is rejected when any value is non-finite or outside [T_lo, T_hi].
    Returns (T_new, hist), T_new None on rejection."""
    d = T_old.dtype.type
    LT0 = (h.mop(T0, 0) - T0) / d(h.dt)
    rhs = mask * ((T_old + d(dt) * src) - (T0 + d(dt) * LT0))
    theta0 = mask * (T_old - T0)
    theta, hist = h.solve(rhs, x0=theta0, max_cycles=max_cycles, tol=tol)
    theta = theta * mask
    cand = T0 + theta
    if not np.isfinite(cand).all() or cand.min() < d(T_lo) or cand.max() > d(T_hi):
        return None, hist
    return reconstruct(T0, theta, h.ax, h.levels[0]["out_idx"]), hist


def chebyshev_smooth(h: Hierarchy, level, x, b, n_steps=2, lam_min_ratio=0.1):
    """chebyshev_smooth (mg_pcg.py:195-247) on level ``level`` of ``h`` (dtype of x);
    returns the smoothed x."""
    lv = h.levels[level]
    d_ = x.dtype.type
    lam_max = float((lv["co"][6] * d_(h.dt) + d_(1.0)).max())
    lam_min = lam_min_ratio * lam_max
    d = (lam_max + lam_min) / 2.0
    c = (lam_max - lam_min) / 2.0
    m, inv = lv["mask"], lv["invD"]
    r = (b - h.mop(x, level)) * m
    p = inv * r
    x_new = x + p
    if n_steps == 1:
        return x_new
    rho_prev = 1.0
    for _ in range(1, n_steps):
        rho_new = 1.0 / (2.0 * d / c - rho_prev)
        r = (b - h.mop(x_new, level)) * m
        p = d_(rho_new) * (d_(2.0 * d / c) * (inv * r) + d_(rho_prev) * p)
        x_new = x_new + p
        rho_prev = rho_new
    return x_new * m


def fcg_mg_fp64(h: Hierarchy, b, tol=1e-8, max_it=100, truncation=1):
    """Flexible CG (Notay 2000, truncation 1) with FP64 vectors and residual;
    the preconditioner (one FP32 V-cycle of ``h``) sees the residual cast to
    the hierarchy dtype. The true residual b - A x in FP64 judges convergence.
    Reference algorithm for the build's FP64-outer solver (no counterpart in
    the reference)."""
    m64 = h.levels[0]["mask"].astype(np.float64)
    co64 = tuple(c.astype(np.float64) for c in h.levels[0]["co"])
    A = lambda v: mop(v, co64, h.dt, h.ax, h.levels[0]["out_idx"]) * m64  # noqa: E731
    b = b.astype(np.float64) * m64
    x = np.zeros_like(b)
    r = b.copy()
    bn = float(np.linalg.norm(b)) + 1e-300
    P, AP, PAP = [], [], []
    hist = [float(np.linalg.norm(r)) / bn]
    it = 0
    while hist[-1] >= tol and it < max_it:
        z = h.precond(r.astype(h.levels[0]["co"][0].dtype)).astype(np.float64)
        p = z.copy()
        for pj, apj, papj in zip(P[-truncation:], AP[-truncation:], PAP[-truncation:]):
            p -= (float(np.dot(z.ravel(), apj.ravel())) / papj) * pj
        ap = A(p)
        pap = float(np.dot(p.ravel(), ap.ravel()))
        alpha = float(np.dot(p.ravel(), r.ravel())) / pap
        x += alpha * p
        r -= alpha * ap
        P.append(p); AP.append(ap); PAP.append(pap)
        P, AP, PAP = P[-truncation:], AP[-truncation:], PAP[-truncation:]
        it += 1
        hist.append(float(np.linalg.norm(b - A(x))) / bn)
    return x, {"it": it, "hist": hist, "converged": hist[-1] < tol}


def dense_operator(co, dt, ax, out_idx, mask=None):
    """Dense matrix of Lo = T + dt L(T) (rows of masked nodes become identity-free
    zero rows when mask is given) — brute force for tiny grids."""
    shape = co[0].shape
    n = int(np.prod(shape))
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = mop(e.reshape(shape), tuple(c.astype(np.float64) for c in co), dt, ax, out_idx).ravel()
    return A
