"""Krylov solvers preconditioned by one multigrid V-cycle (drop-in for solver/mg_pcg.py).

``pcg_mg(mg_solver, b, x0, tol, max_it)`` keeps the reference's signature and
result dictionary (mg_pcg.py:33-97). The iteration is flexible CG (Notay
2000, one previous direction) with every outer vector, dot product and update
in FP64 on the GPU and the V-cycle in FP32: the V-cycle with averaging
restriction, trilinear prolongation and 0.75 damping is not symmetric, so
plain CG can stall before 1e-8, while flexible CG is CG whenever the
preconditioner is symmetric. Convergence is judged on the true FP64 residual
||b - A x|| / ||b|| (checked whenever the recursive residual drops below tol).

All scalars (p.Ap, p.r, alpha, beta, r.r) stay in device memory. The stopping
test runs on the device (fc_solver_stop_check) and the host reads its result one
iteration late from pinned memory, so it queues iteration k+1 while the GPU still
runs iteration k; once the test has stopped the solve, the kernels of the queued
iteration leave x, r and the directions unchanged (FCG and GCR).

Also here: ``bicgstab_mg`` (the reference's BiCGSTAB branch, mg_pcg.py:103-175) and
``patch_mg_pcg``'s switch to it; flexible GCR(m) (``gcr_fp64``) and restarted flexible
GMRES(m) (``fgmres_fp64``), minimal-residual methods for the non-symmetric systems.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import torch

from .. import _lib
from . import mg_cuda_ext  # noqa: F401
from .nlevel_mg_cuda import SKIP_RUN

S_PAP, S_PR, S_RR, S_BB = 0, 1, 2, 3

_lib.register({
    "fc_solver_stop_check": (_lib.I32, [_lib.P, _lib.I32, _lib.I32, _lib.F64, _lib.F64, _lib.I32, _lib.P]),
    "fc_mg_stop_slot": (_lib.I32, []),
    "fc_gcr_stop_slot": (_lib.I32, []),
}, unchecked={"fc_mg_stop_slot", "fc_gcr_stop_slot"})


class _Lookahead:
    """Host side of the device stopping test (fc_solver_stop_check): after each queued
    iteration its r.r and stop flag (scalar slots rr_slot..stop_slot) are copied to
    pinned memory behind its kernels, and the host reads them one iteration later, so
    it never waits for the GPU to drain before it queues the next iteration."""

    def __init__(self, scal, rr_slot, stop_slot, bnorm, tol, stream):
        self.scal, self.rr_slot, self.stop_slot = scal, rr_slot, stop_slot
        self.bnorm, self.tol, self.stream = bnorm, tol, stream
        self.host = torch.zeros((2, stop_slot - rr_slot + 1), dtype=torch.float64, pin_memory=True)
        self.events = [torch.cuda.Event(), torch.cuda.Event()]

    def _check(self, clear):
        _lib.call("fc_solver_stop_check", self.scal.data_ptr(), self.rr_slot, self.stop_slot, self.bnorm, self.tol,
                  clear, self.stream)

    def clear(self):
        self._check(1)

    def push(self, j):
        self._check(0)
        self.host[j & 1].copy_(self.scal[self.rr_slot:self.stop_slot + 1], non_blocking=True)
        self.events[j & 1].record()

    def pull(self, j):
        self.events[j & 1].synchronize()
        return float(self.host[j & 1, 0]), float(self.host[j & 1, -1]) != 0.0


# Grids from this many cells up queue each iteration before reading the previous one's
# stopping test (at 128^3: FCG 34 ms instead of 41 ms, GCR(6) 23 ms instead of 25 ms;
# equal at 256^3). Below it the host's launches take longer than the GPU's work, and
# the test's extra launch and copy cost more than the wait they save.
LOOKAHEAD_MIN_CELLS = 1 << 20


def _use_lookahead(lookahead, n, on_iter, check_every):
    if on_iter is not None or check_every != 1:
        return False  # the callback and sparse checks need the loop that waits after every step
    return n >= LOOKAHEAD_MIN_CELLS if lookahead is None else bool(lookahead)


def _clear_stop(scal, rr_slot, stop_slot, stream):
    _lib.call("fc_solver_stop_check", scal.data_ptr(), rr_slot, stop_slot, 1.0, 0.0, 1, stream)


def _iterate_ahead(step, look, bnorm, tol, max_it, true_residual, hist):
    """step(j) for j = 0, 1, ..., each queued one iteration ahead of its stopping test.
    The history, the iteration count and the iterate equal those of a loop that tests
    after every step: the step queued past the stopping point runs with the device
    flag set and leaves x, r and the directions unchanged. Returns (iterations,
    converged)."""
    step(0)
    look.push(0)
    queued, j = 1, 0
    while True:
        if queued < max_it:
            step(queued)
            look.push(queued)
            queued += 1
        rr, stop = look.pull(j)
        it = j + 1
        rec = rr ** 0.5 / bnorm
        if stop:
            if not (rec < float("inf")):
                hist.append(rec)
                return it, False
            rr_t, _ = true_residual()  # the FP64 true residual decides convergence
            hist.append(rr_t ** 0.5 / bnorm)
            if hist[-1] < tol:
                return it, True
            if it >= max_it:
                return it, False
            look.clear()
            queued = it  # the step queued behind this one changed nothing: queue it again
        else:
            hist.append(rec)
            if it >= max_it:
                return it, False
        j = it


class FCGWorkspace:
    def __init__(self, shape, device):
        n = int(torch.Size(shape).numel())
        f64 = dict(device=device, dtype=torch.float64)
        self.x = torch.zeros(n, **f64)
        self.r = torch.zeros(n, **f64)
        self.b = torch.zeros(n, **f64)
        self.p = torch.zeros(n, **f64)
        self.Ap = torch.zeros(n, **f64)
        self.r32 = torch.zeros(n, device=device, dtype=torch.float32)
        self.z32 = torch.zeros(n, device=device, dtype=torch.float32)
        self.partial = torch.empty(3 * int(_lib.call("fc_mg_partials_count", n)), **f64)
        self.scal = torch.zeros(int(_lib.call("fc_mg_scalar_slots")), **f64)
        self.n = n


def fcg_fp64(mg_solver, b: torch.Tensor, x0: Optional[torch.Tensor] = None, tol: float = 1e-8, max_it: int = 200,
             precond=None, ws: Optional[FCGWorkspace] = None, check_every: int = 1,
             on_iter=None, lookahead: Optional[bool] = None) -> Tuple[torch.Tensor, dict]:
    """Flexible CG on A x = b (A = I + dt L of mg_solver's finest level, masked rows
    excluded). ``precond(r32, z32)`` defaults to one V-cycle of mg_solver; it may
    read ||r||^2 from ``ws.scal[S_RR]`` (device FP64). ``on_iter(it, ws)`` is
    called before each preconditioner application (it = 0 for the initial
    residual). ``lookahead`` (default: from LOOKAHEAD_MIN_CELLS cells up, never with
    on_iter or check_every > 1) queues each iteration before the previous one's
    stopping test is read, with the same result."""
    lv0 = mg_solver._levels[0]
    shape = lv0["shape"]
    nx, ny, nz = shape
    dev = mg_solver.device
    if ws is None:
        ws = FCGWorkspace(shape, dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    if precond is None:
        def precond(r32, z32):
            mg_solver.precondition(r32, out=z32, masked_input=True)
    ws.b.copy_(b.reshape(-1).to(torch.float64))
    if x0 is None:
        ws.x.zero_()
    else:  # Dirichlet rows start at 0 (mg_pcg.py:50-51): the stencil reads them
        ws.x.copy_(x0.reshape(-1).to(torch.float64))
        ws.x.mul_(lv0["mask_f"].reshape(-1))
    dt = float(mg_solver._dt_cache)
    ax, oi = mg_solver.ax, lv0["out_idx"]
    mask = lv0["mask_f"]
    rows = _lib.ptr(lv0.get("rows"))  # runs of SKIP_RUN cells without unknowns, skipped (NULL: none)

    def true_residual():
        _lib.call("fc_mg_true_residual", ws.x.data_ptr(), ws.b.data_ptr(), ws.r.data_ptr(), ws.r32.data_ptr(),
                  mask.data_ptr(), lv0["coef"], dt, nx, ny, nz, ax, oi, ws.partial.data_ptr(), ws.scal.data_ptr(),
                  stream)
        s = ws.scal[S_RR:S_BB + 1].cpu()
        return float(s[0]), float(s[1])

    def step(j):
        precond(ws.r32, ws.z32)
        _lib.call("fc_mg_new_direction", ws.p.data_ptr(), ws.z32.data_ptr(), ws.Ap.data_ptr(), ws.n, int(j == 0),
                  rows, SKIP_RUN, ws.partial.data_ptr(), ws.scal.data_ptr(), stream)
        _lib.call("fc_mg_apply_A64", ws.p.data_ptr(), ws.Ap.data_ptr(), ws.r.data_ptr(), mask.data_ptr(),
                  lv0["coef"], dt, nx, ny, nz, ax, oi, ws.partial.data_ptr(), ws.scal.data_ptr(), stream)
        _lib.call("fc_mg_update_xr", ws.x.data_ptr(), ws.r.data_ptr(), ws.r32.data_ptr(), ws.p.data_ptr(),
                  ws.Ap.data_ptr(), ws.n, rows, SKIP_RUN, ws.partial.data_ptr(), ws.scal.data_ptr(), stream)

    S_STOP = _stop_slot("mg")
    _clear_stop(ws.scal, S_RR, S_STOP, stream)  # the workspace may hold a finished solve's flag
    rr, bb = true_residual()
    bnorm = bb ** 0.5 + 1e-300
    hist = [rr ** 0.5 / bnorm]
    it = 0
    converged = hist[-1] < tol
    if not converged and max_it > 0 and _use_lookahead(lookahead, ws.n, on_iter, check_every):
        look = _Lookahead(ws.scal, S_RR, S_STOP, bnorm, tol, stream)
        it, converged = _iterate_ahead(step, look, bnorm, tol, max_it, true_residual, hist)
        max_it = it  # done: the loop below tests after every step for on_iter / check_every
    while not converged and it < max_it:
        if on_iter is not None:
            on_iter(it, ws)
        step(it)
        it += 1
        if it % check_every and it < max_it:
            continue
        rec = float(ws.scal[S_RR].item()) ** 0.5 / bnorm
        if not (rec < float("inf")):
            hist.append(rec)
            break
        if rec < tol:
            rr, _ = true_residual()  # the FP64 true residual decides convergence
            hist.append(rr ** 0.5 / bnorm)
            converged = hist[-1] < tol
        else:
            hist.append(rec)
    x = ws.x.view(shape).to(torch.float32) * mask
    return x, {"converged": converged, "it": it, "res_rel": hist[-1], "hist": hist, "x64": ws.x.view(shape)}


def pcg_mg(mg_solver, b, x0=None, tol: float = 1e-3, max_it: int = 15):
    """mg_pcg.py:33-97 surface: returns (x, {"converged", "it", "res_rel"})."""
    x, info = fcg_fp64(mg_solver, b, x0=x0, tol=tol, max_it=max_it)
    return x, {"converged": info["converged"], "it": info["it"], "res_rel": info["res_rel"]}


_lib.register({
    "fc_mg_cheb_partials_count": (_lib.I32, [_lib.I64]),
    "fc_mg_cheb_lambda_max": (_lib.I32, [_lib.P, _lib.I64, _lib.F64, _lib.P, _lib.P, _lib.P]),
    "fc_mg_cheb_first": (_lib.I32, [_lib.P, _lib.P, _lib.P, _lib.P, _lib.P, _lib.P, _lib.P, _lib.I64, _lib.P]),
    "fc_mg_cheb_step": (_lib.I32, [_lib.P, _lib.P, _lib.P, _lib.P, _lib.P, _lib.P, _lib.F64, _lib.F64, _lib.F64,
                                   _lib.P, _lib.I64, _lib.P]),
}, unchecked={"fc_mg_cheb_partials_count"})


def chebyshev_smooth(mg_solver, level: int, n_steps: int = 2, lam_min_ratio: float = 0.1) -> None:
    """Chebyshev smoothing of levels[level]['x'] against levels[level]['b_buf'], in
    place (mg_pcg.py:195-247): lambda_max = max(diagL dt + 1), lambda_min =
    lam_min_ratio lambda_max, d and c the centre and half-width of that interval;
    x_1 = x + invD r, then n_steps - 1 recurrences p = rho (2 d / c invD r +
    rho_prev p) with rho = 1 / (2 d / c - rho_prev); the result is masked (not for
    n_steps = 1, as in the reference). Every array operation runs in HIP kernels."""
    lv = mg_solver._levels[level]
    n = int(lv["x"].numel())
    stream = mg_solver._stream()
    if "cheb_p" not in lv:
        lv["cheb_p"] = torch.empty_like(lv["x"])
        lv["cheb_x"] = torch.empty_like(lv["x"])
        lv["cheb_partial"] = torch.empty(int(_lib.call("fc_mg_cheb_partials_count", n)) + 1, dtype=torch.float32,
                                         device=lv["x"].device)
    part = lv["cheb_partial"]
    _lib.call("fc_mg_cheb_lambda_max", lv["diagL"].data_ptr(), n, float(mg_solver._dt_cache), part[1:].data_ptr(),
              part.data_ptr(), stream)
    lam_max = float(part[0].item())
    lam_min = lam_min_ratio * lam_max
    d = (lam_max + lam_min) / 2.0
    c = (lam_max - lam_min) / 2.0
    x, b, p, x_new = lv["x"], lv["b_buf"], lv["cheb_p"], lv["cheb_x"]
    inv, mask = lv["inv_diagM"], lv["mask_f"]
    mg_solver._Mop(x, level)
    _lib.call("fc_mg_cheb_first", x.data_ptr(), b.data_ptr(), lv["Lo"].data_ptr(), inv.data_ptr(), mask.data_ptr(),
              p.data_ptr(), x_new.data_ptr(), n, stream)
    if n_steps == 1:
        x.copy_(x_new)
        return
    rho_prev = 1.0
    for k in range(1, n_steps):
        rho_new = 1.0 / (2.0 * d / c - rho_prev)
        mg_solver._Mop(x_new, level)
        _lib.call("fc_mg_cheb_step", x_new.data_ptr(), b.data_ptr(), lv["Lo"].data_ptr(), inv.data_ptr(),
                  mask.data_ptr(), p.data_ptr(), rho_new, 2.0 * d / c, rho_prev,
                  x.data_ptr() if k == n_steps - 1 else None, n, stream)
        rho_prev = rho_new


def patch_mg_pcg(mg_solver, pe_threshold: float = 1.5, nonsymmetric_route: Optional[str] = None):
    """Route mg_solver.solve through PCG or BiCGSTAB (mg_pcg.py:250-275): with the
    reference's heuristic pe = 0.1 * max(diagL) of the finest level, above
    ``pe_threshold`` bicgstab_mg, else pcg_mg.

    The heuristic looks at diagL only, so a non-symmetric system (advection, or a
    varying rho*c_p) with a small diagL goes to (flexible) CG, which can diverge on it:
    the 192x192x384 heat sink with advection reaches a relative residual of 4e24.
    ``nonsymmetric_route`` = "gcr" or "fgmres" (default None: the reference's routing)
    sends every system whose finest level keeps per-cell coefficients (not the
    symmetric face form) to gcr_fp64 (6 directions) or fgmres_fp64 (16), which
    minimise the residual. The info dict names the solver that ran ("route")."""
    import types

    if nonsymmetric_route not in (None, "gcr", "fgmres"):
        raise ValueError("nonsymmetric_route must be None, 'gcr' or 'fgmres'")

    def _solve_pcg(self, b, x0=None, max_cycles=15, tol=1e-3):
        lv0 = self._levels[0]
        if nonsymmetric_route is not None and lv0.get("faces") is None:
            if nonsymmetric_route == "gcr":
                x, info = gcr_fp64(self, b, x0=x0, tol=tol, max_it=max_cycles, m=6)
            else:
                x, info = fgmres_fp64(self, b, x0=x0, tol=tol, max_it=max_cycles, m=16)
            info["route"] = nonsymmetric_route
            return x, info
        pe = float(lv0["diagL"].max()) * 0.1
        if pe > pe_threshold:
            x, info = bicgstab_mg(self, b, x0, tol=tol, max_it=max_cycles)
            info["route"] = "bicgstab"
        else:
            x, info = pcg_mg(self, b, x0, tol=tol, max_it=max_cycles)
            info["route"] = "pcg"
        return x, info

    mg_solver.solve = types.MethodType(_solve_pcg, mg_solver)
    return mg_solver


# ----------------------------------------------------------- flexible GCR(m)
_lib.register({
    "fc_gcr_scalar_slots": (_lib.I32, []),
    "fc_gcr_max_directions": (_lib.I32, []),
    "fc_gcr_partials_count": (_lib.I32, [_lib.I64]),
    "fc_gcr_step": (_lib.I32, [_lib.P, _lib.P, _lib.P, _lib.P, _lib.P, _lib.P, _lib.P, _lib.I32, _lib.I32, _lib.P,
                                _lib.P, _lib.P, _lib.P, _lib.P, _lib.P]),
}, unchecked={"fc_gcr_scalar_slots", "fc_gcr_max_directions", "fc_gcr_partials_count"})

K_RR = 2
K_WWHIST0 = 16


def _stop_slot(which: str) -> int:
    """Scalar slot of the device stopping flag of the FCG ("mg") or GCR ("gcr") workspace."""
    return int(_lib.call(f"fc_{which}_stop_slot"))


class GCRStencil(ctypes.Structure):
    _fields_ = [("aE", ctypes.c_void_p), ("aW", ctypes.c_void_p), ("aN", ctypes.c_void_p), ("aS", ctypes.c_void_p),
                ("aU", ctypes.c_void_p), ("aD", ctypes.c_void_p), ("diagL", ctypes.c_void_p),
                ("mask", ctypes.c_void_p), ("dt", ctypes.c_double), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("nz", ctypes.c_int), ("ax", ctypes.c_int), ("out_idx", ctypes.c_int), ("fx", ctypes.c_void_p),
                ("fy", ctypes.c_void_p), ("fz", ctypes.c_void_p)]


class GCRWorkspace:
    def __init__(self, shape, device, m: int = 6):
        mmax = int(_lib.call("fc_gcr_max_directions"))
        if not 0 <= m < mmax:
            raise ValueError(f"GCR truncation m must be in [0, {mmax - 1}]")
        n = int(torch.Size(shape).numel())
        f64 = dict(device=device, dtype=torch.float64)
        self.m, self.n = m, n
        self.fcg = FCGWorkspace(shape, device)  # x, r, b, r32, z32 and the true-residual reductions
        self.P = [torch.zeros(n, **f64) for _ in range(m + 1)]
        self.W = [torch.zeros(n, **f64) for _ in range(m + 1)]
        self.partial = torch.empty(int(_lib.call("fc_gcr_partials_count", n)), **f64)
        self.scal = torch.zeros(int(_lib.call("fc_gcr_scalar_slots")), **f64)


def gcr_fp64(mg_solver, b: torch.Tensor, x0: Optional[torch.Tensor] = None, tol: float = 1e-8, max_it: int = 300,
             precond=None, m: int = 6, ws: Optional[GCRWorkspace] = None, on_iter=None,
             check_every: int = 1, lookahead: Optional[bool] = None) -> Tuple[torch.Tensor, dict]:
    """Flexible GCR(m) on A x = b with FP64 vectors and a variable preconditioner
    ``precond(r32, z32, ws)`` (default: one V-cycle). The residual norm never
    increases; convergence is judged on the FP64 true residual. ``on_iter(it,
    rel_res)`` may raise to stop the iteration early. ``lookahead`` as in fcg_fp64."""
    lv0 = mg_solver._levels[0]
    shape = lv0["shape"]
    nx, ny, nz = shape
    dev = mg_solver.device
    if ws is None:
        ws = GCRWorkspace(shape, dev, m)
    fw = ws.fcg
    stream = torch.cuda.current_stream(dev).cuda_stream
    if precond is None:
        def precond(r32, z32, _ws):
            mg_solver.precondition(r32, out=z32, masked_input=True)
    fw.b.copy_(b.reshape(-1).to(torch.float64))
    if x0 is None:
        fw.x.zero_()
    else:  # Dirichlet rows start at 0 (mg_pcg.py:50-51): the stencil reads them
        fw.x.copy_(x0.reshape(-1).to(torch.float64))
        fw.x.mul_(lv0["mask_f"].reshape(-1))
    dt = float(mg_solver._dt_cache)
    mask = lv0["mask_f"]
    faces = lv0.get("faces")
    st = GCRStencil(*[lv0[k].data_ptr() for k in ("aE", "aW", "aN", "aS", "aU", "aD", "diagL")], mask.data_ptr(),
                    dt, nx, ny, nz, int(mg_solver.ax), int(lv0["out_idx"]),
                    *([f.data_ptr() for f in faces] if faces is not None else [None] * 3))

    def true_residual():
        _lib.call("fc_mg_true_residual", fw.x.data_ptr(), fw.b.data_ptr(), fw.r.data_ptr(), fw.r32.data_ptr(),
                  mask.data_ptr(), lv0["coef"], dt, nx, ny, nz, mg_solver.ax, lv0["out_idx"], fw.partial.data_ptr(),
                  fw.scal.data_ptr(), stream)
        s = fw.scal[S_RR:S_BB + 1].cpu()
        return float(s[0]), float(s[1])

    W_ptrs = (ctypes.c_void_p * max(m, 1))()
    P_ptrs = (ctypes.c_void_p * max(m, 1))()
    slots = (ctypes.c_int * max(m, 1))()

    def step(i):
        precond(fw.r32, fw.z32, ws)
        cur = i % (m + 1)
        k = min(i, m)
        for j in range(k):
            src = (i - 1 - j) % (m + 1)
            W_ptrs[j], P_ptrs[j], slots[j] = ws.W[src].data_ptr(), ws.P[src].data_ptr(), K_WWHIST0 + src
        _lib.call("fc_gcr_step", fw.z32.data_ptr(), ctypes.byref(st), ws.P[cur].data_ptr(), ws.W[cur].data_ptr(),
                  W_ptrs, P_ptrs, slots, k, K_WWHIST0 + cur, fw.x.data_ptr(), fw.r.data_ptr(), fw.r32.data_ptr(),
                  ws.partial.data_ptr(), ws.scal.data_ptr(), stream)

    K_STOP = _stop_slot("gcr")
    _clear_stop(ws.scal, K_RR, K_STOP, stream)  # the workspace may hold a finished solve's flag
    rr, bb = true_residual()
    bnorm = bb ** 0.5 + 1e-300
    hist = [rr ** 0.5 / bnorm]
    converged = hist[-1] < tol
    it = 0
    stopped = False
    if not converged and max_it > 0 and _use_lookahead(lookahead, ws.n, on_iter, check_every):
        look = _Lookahead(ws.scal, K_RR, K_STOP, bnorm, tol, stream)
        it, converged = _iterate_ahead(step, look, bnorm, tol, max_it, true_residual, hist)
        max_it = it  # done: the loop below tests after every step for on_iter / check_every
    while not converged and it < max_it:
        if on_iter is not None:
            try:
                on_iter(it, hist[-1])
            except StopIteration:
                stopped = True
                break
        step(it)
        it += 1
        if it % check_every and it < max_it:
            continue
        rec = float(ws.scal[K_RR].item()) ** 0.5 / bnorm
        if not (rec < float("inf")):
            hist.append(rec)
            break
        if rec < tol:
            rr, _ = true_residual()
            hist.append(rr ** 0.5 / bnorm)
            converged = hist[-1] < tol
        else:
            hist.append(rec)
    x = fw.x.view(shape).to(torch.float32) * mask
    return x, {"converged": converged, "it": it, "res_rel": hist[-1], "hist": hist, "x64": fw.x.view(shape),
               "stopped": stopped}


# ---------------------------------------------------- restarted flexible GMRES(m)
_lib.register({
    "fc_fgmres_max_basis": (_lib.I32, []),
    "fc_fgmres_scalar_slots": (_lib.I32, []),
    "fc_fgmres_dot_slot": (_lib.I32, [_lib.I32]),
    "fc_fgmres_norm_slot": (_lib.I32, [_lib.I32]),
    "fc_fgmres_partials_count": (_lib.I32, [_lib.I64]),
    "fc_fgmres_arnoldi": (_lib.I32, [_lib.P, _lib.P, _lib.P, _lib.I32, _lib.P, _lib.P, _lib.P, _lib.P]),
    "fc_fgmres_update_x": (_lib.I32, [_lib.P, _lib.P, _lib.P, _lib.I32, _lib.I64, _lib.P]),
    "fc_fgmres_givens_slots": (_lib.I32, []),
    "fc_fgmres_givens_layout": (_lib.I32, [_lib.P, _lib.P, _lib.P, _lib.P]),
    "fc_fgmres_givens": (_lib.I32, [_lib.P, _lib.P, _lib.I32, _lib.F64, _lib.P]),
}, unchecked={"fc_fgmres_max_basis", "fc_fgmres_scalar_slots", "fc_fgmres_dot_slot", "fc_fgmres_norm_slot",
              "fc_fgmres_partials_count", "fc_fgmres_givens_slots"})


class FGMRESWorkspace:
    """Arnoldi basis V_1..V_m (FP64; V_0 is the FP64 true-residual buffer), the m
    preconditioned vectors Z_j in FP32 (the V-cycle's own output, stored as is), the
    FP32 copy of the newest basis vector and the device scalars."""

    def __init__(self, shape, device, m: int = 16):
        mmax = int(_lib.call("fc_fgmres_max_basis"))
        if not 1 <= m <= mmax:
            raise ValueError(f"FGMRES restart length m must be in [1, {mmax}]")
        n = int(torch.Size(shape).numel())
        self.m, self.n = m, n
        self.fcg = FCGWorkspace(shape, device)  # x, b, r (= V_0), r32 and the true-residual reductions
        self.V = [torch.zeros(n, device=device, dtype=torch.float64) for _ in range(m)]
        self.Z = [torch.zeros(n, device=device, dtype=torch.float32) for _ in range(m)]
        self.v32 = torch.zeros(n, device=device, dtype=torch.float32)
        self.partial = torch.empty(int(_lib.call("fc_fgmres_partials_count", n)), device=device, dtype=torch.float64)
        self.scal = torch.zeros(int(_lib.call("fc_fgmres_scalar_slots")), device=device, dtype=torch.float64)
        self.dot_slot = [int(_lib.call("fc_fgmres_dot_slot", i)) for i in range(m + 1)]
        self.norm_slot = [int(_lib.call("fc_fgmres_norm_slot", i)) for i in range(m + 1)]
        lay = [ctypes.c_int() for _ in range(4)]
        _lib.call("fc_fgmres_givens_layout", *[ctypes.byref(v) for v in lay])
        self.g_r, self.g_g, self.g_est, self.g_ld = (v.value for v in lay)
        self.gstate = torch.zeros(int(_lib.call("fc_fgmres_givens_slots")), device=device, dtype=torch.float64)
        # residual estimates come back through pinned memory, one step behind
        self.est_host = torch.zeros(m, dtype=torch.float64, pin_memory=True)
        self.events = [torch.cuda.Event() for _ in range(m)]


def fgmres_fp64(mg_solver, b: torch.Tensor, x0: Optional[torch.Tensor] = None, tol: float = 1e-8,
                max_it: int = 300, m: int = 16, precond=None, ws: Optional[FGMRESWorkspace] = None
                ) -> Tuple[torch.Tensor, dict]:
    """Restarted flexible GMRES(m) (Saad 1993) on A x = b with FP64 vectors and a
    variable preconditioner ``precond(v32, z32)`` (default: one V-cycle).

    Each Arnoldi step is two passes over the grid (fc_fgmres_arnoldi): w = A z with
    the dot products against the basis, then classical Gram-Schmidt in place with
    the FP32 copy for the next preconditioner call. The basis is kept unnormalised;
    a one-thread kernel (fc_fgmres_givens) forms the normalised Hessenberg column
    (h_ij = w.V_i / (|V_j| |V_i|), h_j+1,j = |V_j+1| / |V_j|) and applies the Givens
    rotations in FP64 on the device. The host reads each residual estimate one step
    late, from pinned memory, while the next step is already queued, so the GPU does
    not wait for it; convergence seen at step j ends the cycle after step j + 1, whose
    direction is used too. At each restart x += sum_j (y_j / |V_j|) Z_j and the FP64
    true residual decides convergence and starts the next cycle."""
    import math

    lv0 = mg_solver._levels[0]
    shape = lv0["shape"]
    nx, ny, nz = shape
    dev = mg_solver.device
    if ws is None or ws.m != m:
        ws = FGMRESWorkspace(shape, dev, m)
    fw = ws.fcg
    stream = torch.cuda.current_stream(dev).cuda_stream
    if precond is None:
        def precond(v32, z32):
            mg_solver.precondition(v32, out=z32, masked_input=True)
    fw.b.copy_(b.reshape(-1).to(torch.float64))
    if x0 is None:
        fw.x.zero_()
    else:  # Dirichlet rows start at 0 (mg_pcg.py:50-51): the stencil reads them
        fw.x.copy_(x0.reshape(-1).to(torch.float64))
        fw.x.mul_(lv0["mask_f"].reshape(-1))
    dt = float(mg_solver._dt_cache)
    mask = lv0["mask_f"]
    faces = lv0.get("faces")
    st = GCRStencil(*[lv0[k].data_ptr() for k in ("aE", "aW", "aN", "aS", "aU", "aD", "diagL")], mask.data_ptr(),
                    dt, nx, ny, nz, int(mg_solver.ax), int(lv0["out_idx"]),
                    *([f.data_ptr() for f in faces] if faces is not None else [None] * 3))

    def true_residual():
        _lib.call("fc_mg_true_residual", fw.x.data_ptr(), fw.b.data_ptr(), fw.r.data_ptr(), fw.r32.data_ptr(),
                  mask.data_ptr(), lv0["coef"], dt, nx, ny, nz, mg_solver.ax, lv0["out_idx"], fw.partial.data_ptr(),
                  fw.scal.data_ptr(), stream)
        s = fw.scal[S_RR:S_BB + 1].cpu()
        return float(s[0]), float(s[1])

    Vp = (ctypes.c_void_p * (m + 1))(fw.r.data_ptr(), *[v.data_ptr() for v in ws.V])
    Zp = (ctypes.c_void_p * m)(*[z.data_ptr() for z in ws.Z])
    rr, bb = true_residual()
    bnorm = bb ** 0.5 + 1e-300
    hist = [rr ** 0.5 / bnorm]
    converged = hist[-1] < tol
    it = 0
    while not converged and it < max_it and rr > 0.0:
        ws.scal[ws.norm_slot[0]:ws.norm_slot[0] + 1].copy_(fw.scal[S_RR:S_RR + 1])  # |V_0|^2 on the device
        vin, k = fw.r32, 0
        for j in range(m):
            precond(vin, ws.Z[j])
            _lib.call("fc_fgmres_arnoldi", ws.Z[j].data_ptr(), ctypes.byref(st), Vp, j, ws.v32.data_ptr(),
                      ws.partial.data_ptr(), ws.scal.data_ptr(), stream)
            _lib.call("fc_fgmres_givens", ws.scal.data_ptr(), ws.gstate.data_ptr(), j, bnorm, stream)
            ws.est_host[j:j + 1].copy_(ws.gstate[ws.g_est:ws.g_est + 1], non_blocking=True)
            ws.events[j].record()
            it += 1
            k = j + 1
            if j >= 1:  # the estimate of the previous step; this step is already queued
                ws.events[j - 1].synchronize()
                e = float(ws.est_host[j - 1])
                hist.append(e)
                if not e >= 0.5 * tol:  # converged (or not finite): stop after this step
                    break
            if it >= max_it:
                break
            vin = ws.v32
        ws.events[k - 1].synchronize()
        hist.append(float(ws.est_host[k - 1]))
        gst = ws.gstate.cpu().tolist()
        nv = [float(v) for v in ws.scal[ws.norm_slot[0]:ws.norm_slot[0] + k].cpu()]
        # an exact breakdown (V_i = 0, or an empty pivot) ends the basis at column i:
        # that column and any queued after it do not enter the update
        k_used = next((i for i in range(k) if not (nv[i] > 0.0 and gst[ws.g_r + i * ws.g_ld + i] != 0.0)), k)
        it -= k - k_used
        k = k_used
        if k == 0:
            break
        y = [0.0] * k
        for i in range(k - 1, -1, -1):
            rii = gst[ws.g_r + i * ws.g_ld + i]
            acc = gst[ws.g_g + i] - sum(gst[ws.g_r + q * ws.g_ld + i] * y[q] for q in range(i + 1, k))
            y[i] = acc / rii if rii != 0.0 else 0.0
        a = (ctypes.c_double * k)(*[y[i] / nv[i] ** 0.5 for i in range(k)])
        _lib.call("fc_fgmres_update_x", fw.x.data_ptr(), Zp, a, k, ws.n, stream)
        rr, _ = true_residual()
        hist.append(rr ** 0.5 / bnorm)
        converged = hist[-1] < tol
        if not math.isfinite(hist[-1]):
            break
    x = fw.x.view(shape).to(torch.float32) * mask
    return x, {"converged": converged, "it": it, "res_rel": hist[-1], "hist": hist, "x64": fw.x.view(shape)}


# ------------------------------------------------- MG-preconditioned BiCGSTAB
_lib.register({
    "fc_krylov_apply": (_lib.I32, [_lib.P, _lib.P, _lib.P, _lib.I32, _lib.P]),
    "fc_krylov_dots": (_lib.I32, [_lib.P, _lib.P, _lib.I32, _lib.I64, _lib.P, _lib.P, _lib.I32, _lib.P]),
    "fc_bicgstab_direction": (_lib.I32, [_lib.P, _lib.P, _lib.P, _lib.P, _lib.F64, _lib.F64, _lib.P, _lib.I64,
                                         _lib.P]),
    "fc_bicgstab_s": (_lib.I32, [_lib.P, _lib.P, _lib.P, _lib.P, _lib.F64, _lib.P, _lib.I64, _lib.P]),
    "fc_bicgstab_update": (_lib.I32, [_lib.P, _lib.P, _lib.P, _lib.P, _lib.P, _lib.P, _lib.P, _lib.F64, _lib.F64,
                                      _lib.I64, _lib.P]),
})


class BiCGSTABWorkspace:
    """FP64 vectors x, r, r_hat, p, v, s, t and b; FP32 copies of p and s for the
    V-cycle and its outputs p_hat, s_hat; reduction scratch."""

    def __init__(self, shape, device):
        n = int(torch.Size(shape).numel())
        f64 = dict(device=device, dtype=torch.float64)
        f32 = dict(device=device, dtype=torch.float32)
        self.n = n
        for name in ("x", "r", "r_hat", "p", "v", "s", "t", "b"):
            setattr(self, name, torch.zeros(n, **f64))
        for name in ("p32", "s32", "p_hat", "s_hat"):
            setattr(self, name, torch.zeros(n, **f32))
        self.partial = torch.empty(int(_lib.call("fc_gcr_partials_count", n)), **f64)
        self.scal = torch.zeros(int(_lib.call("fc_gcr_scalar_slots")), **f64)
        self.fcg = FCGWorkspace(shape, device)  # the initial residual (true-residual kernel)


def bicgstab_mg(mg_solver, b, x0=None, tol: float = 1e-3, max_it: int = 15,
                ws: Optional[BiCGSTABWorkspace] = None) -> Tuple[torch.Tensor, dict]:
    """bicgstab_mg (mg_pcg.py:103-175): BiCGSTAB with one V-cycle as preconditioner, for
    operators with advection. Same steps, guards and masking as the reference (v = A
    p_hat and t = A s_hat keep the masked rows; s, r and x are masked; the tolerance is
    on the recursive ||r|| / ||b||), with FP64 vectors and dot products on the GPU and
    the V-cycle in FP32. Returns (x, {"converged", "it", "res_rel"}) like the reference,
    plus "hist" and "x64"."""
    lv0 = mg_solver._levels[0]
    shape = lv0["shape"]
    nx, ny, nz = shape
    dev = mg_solver.device
    if ws is None:
        ws = BiCGSTABWorkspace(shape, dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    mask = lv0["mask_f"]
    dt = float(mg_solver._dt_cache)
    faces = lv0.get("faces")
    st = GCRStencil(*[lv0[k].data_ptr() for k in ("aE", "aW", "aN", "aS", "aU", "aD", "diagL")], mask.data_ptr(),
                    dt, nx, ny, nz, int(mg_solver.ax), int(lv0["out_idx"]),
                    *([f.data_ptr() for f in faces] if faces is not None else [None] * 3))
    n = ws.n
    P2 = ctypes.c_void_p * 3

    def dots(*pairs):
        a = P2(*[p[0].data_ptr() for p in pairs])
        c = P2(*[p[1].data_ptr() for p in pairs])
        _lib.call("fc_krylov_dots", a, c, len(pairs), n, ws.partial.data_ptr(), ws.scal.data_ptr(), 0, stream)
        return ws.scal[:len(pairs)].tolist()

    def precond(v32, out):
        mg_solver.precondition(v32, out=out, masked_input=True)

    ws.b.copy_(b.reshape(-1).to(torch.float64))
    if x0 is None:
        ws.x.zero_()
    else:
        ws.x.copy_(x0.reshape(-1).to(torch.float64))
        ws.x.mul_(mask.reshape(-1))
    fw = ws.fcg
    _lib.call("fc_mg_true_residual", ws.x.data_ptr(), ws.b.data_ptr(), ws.r.data_ptr(), fw.r32.data_ptr(),
              mask.data_ptr(), lv0["coef"], dt, nx, ny, nz, mg_solver.ax, lv0["out_idx"], fw.partial.data_ptr(),
              fw.scal.data_ptr(), stream)
    ws.r_hat.copy_(ws.r)
    bnorm = dots((ws.b, ws.b))[0] ** 0.5 + 1e-30
    rho = alpha = omega = 1.0
    ws.v.zero_()
    ws.p.zero_()
    converged, it, hist = False, 0, []
    for it in range(max_it):
        rr, rho_new = dots((ws.r, ws.r), (ws.r_hat, ws.r))
        res = rr ** 0.5 / bnorm
        hist.append(res)
        if res < tol:
            converged = True
            break
        if abs(rho_new) < 1e-30:
            break
        beta = (rho_new / (rho + 1e-30)) * (alpha / (omega + 1e-30))
        _lib.call("fc_bicgstab_direction", ws.p.data_ptr(), ws.r.data_ptr(), ws.v.data_ptr(), mask.data_ptr(), beta,
                  omega, ws.p32.data_ptr(), n, stream)
        precond(ws.p32, ws.p_hat)
        _lib.call("fc_krylov_apply", ws.p_hat.data_ptr(), ctypes.byref(st), ws.v.data_ptr(), 0, stream)
        alpha = rho_new / (dots((ws.r_hat, ws.v))[0] + 1e-30)
        _lib.call("fc_bicgstab_s", ws.s.data_ptr(), ws.r.data_ptr(), ws.v.data_ptr(), mask.data_ptr(), alpha,
                  ws.s32.data_ptr(), n, stream)
        if dots((ws.s, ws.s))[0] ** 0.5 / bnorm < tol:
            _lib.call("fc_bicgstab_update", ws.x.data_ptr(), None, None, None, ws.p_hat.data_ptr(), None,
                      mask.data_ptr(), alpha, 0.0, n, stream)
            converged = True
            break
        precond(ws.s32, ws.s_hat)
        _lib.call("fc_krylov_apply", ws.s_hat.data_ptr(), ctypes.byref(st), ws.t.data_ptr(), 0, stream)
        ts, tt = dots((ws.t, ws.s), (ws.t, ws.t))
        omega = ts / (tt + 1e-30)
        _lib.call("fc_bicgstab_update", ws.x.data_ptr(), ws.r.data_ptr(), ws.s.data_ptr(), ws.t.data_ptr(),
                  ws.p_hat.data_ptr(), ws.s_hat.data_ptr(), mask.data_ptr(), alpha, omega, n, stream)
        rho = rho_new
    res_final = dots((ws.r, ws.r))[0] ** 0.5 / bnorm
    x = ws.x.view(shape).to(torch.float32) * mask
    return x, {"converged": converged, "it": it + 1, "res_rel": res_final, "hist": hist, "x64": ws.x.view(shape)}
