"""Krylov solvers preconditioned by one multigrid V-cycle (drop-in for solver/mg_pcg.py).

``pcg_mg(mg_solver, b, x0, tol, max_it)`` keeps the reference's signature and
result dictionary (mg_pcg.py:33-97). The iteration is flexible CG (Notay
2000, one previous direction) with every outer vector, dot product and update
in FP64 on the GPU and the V-cycle in FP32: the V-cycle with averaging
restriction, trilinear prolongation and 0.75 damping is not symmetric, so
plain CG can stall before 1e-8, while flexible CG is CG whenever the
preconditioner is symmetric. Convergence is judged on the true FP64 residual
||b - A x|| / ||b|| (checked whenever the recursive residual drops below tol).

All scalars (p.Ap, p.r, alpha, beta, r.r) stay in device memory; the host
reads one FP64 value per iteration for the stopping test.
"""
from __future__ import annotations

from typing import Optional, Tuple

import torch

from .. import _lib
from . import mg_cuda_ext  # noqa: F401

S_PAP, S_PR, S_RR, S_BB = 0, 1, 2, 3


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
             on_iter=None) -> Tuple[torch.Tensor, dict]:
    """Flexible CG on A x = b (A = I + dt L of mg_solver's finest level, masked rows
    excluded). ``precond(r32, z32)`` defaults to one V-cycle of mg_solver; it may
    read ||r||^2 from ``ws.scal[S_RR]`` (device FP64). ``on_iter(it, ws)`` is
    called before each preconditioner application (it = 0 for the initial
    residual)."""
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
    else:
        ws.x.copy_(x0.reshape(-1).to(torch.float64))
    dt = float(mg_solver._dt_cache)
    ax, oi = mg_solver.ax, lv0["out_idx"]
    mask = lv0["mask_f"]

    def true_residual():
        _lib.call("fc_mg_true_residual", ws.x.data_ptr(), ws.b.data_ptr(), ws.r.data_ptr(), ws.r32.data_ptr(),
                  mask.data_ptr(), lv0["coef"], dt, nx, ny, nz, ax, oi, ws.partial.data_ptr(), ws.scal.data_ptr(),
                  stream)
        s = ws.scal[S_RR:S_BB + 1].cpu()
        return float(s[0]), float(s[1])

    rr, bb = true_residual()
    bnorm = bb ** 0.5 + 1e-300
    hist = [rr ** 0.5 / bnorm]
    it = 0
    converged = hist[-1] < tol
    while not converged and it < max_it:
        if on_iter is not None:
            on_iter(it, ws)
        precond(ws.r32, ws.z32)
        _lib.call("fc_mg_new_direction", ws.p.data_ptr(), ws.z32.data_ptr(), ws.Ap.data_ptr(), ws.n, int(it == 0),
                  ws.partial.data_ptr(), ws.scal.data_ptr(), stream)
        _lib.call("fc_mg_apply_A64", ws.p.data_ptr(), ws.Ap.data_ptr(), ws.r.data_ptr(), mask.data_ptr(),
                  lv0["coef"], dt, nx, ny, nz, ax, oi, ws.partial.data_ptr(), ws.scal.data_ptr(), stream)
        _lib.call("fc_mg_update_xr", ws.x.data_ptr(), ws.r.data_ptr(), ws.r32.data_ptr(), ws.p.data_ptr(),
                  ws.Ap.data_ptr(), ws.n, ws.partial.data_ptr(), ws.scal.data_ptr(), stream)
        it += 1
        if it % check_every and it < max_it:
            continue
        rec = float(ws.scal[S_RR].item()) ** 0.5 / bnorm
        if not (rec < float("inf")):
            hist.append(rec)
            break
        if rec < tol:
            rr, _ = true_residual()  # the FP64 true residual is the judge
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


def patch_mg_pcg(mg_solver, pe_threshold: float = 1.5):
    """Route mg_solver.solve through pcg_mg (mg_pcg.py:250-275; the BiCGSTAB
    branch of the reference is not part of this build)."""
    import types

    def _solve_pcg(self, b, x0=None, max_cycles=15, tol=1e-3):
        return pcg_mg(self, b, x0, tol=tol, max_it=max_cycles)

    mg_solver.solve = types.MethodType(_solve_pcg, mg_solver)
    return mg_solver
