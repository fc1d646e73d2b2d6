"""Drop-in replacement of solver/thermal_cuda_ext.py on MI355X.

``THERMAL_CUDA`` exposes thermal_rhs, thermal_reconstruct, convergence_metrics
and physics_check with the reference's arguments and results
(thermal_kernels.cu:198-263); ``_temperature_step_mg_fused`` and
``patch_thermal_fused`` follow thermal_cuda_ext.py:55-234, with the RHS built
by the fused kernel and every reduction on the GPU.
"""
from __future__ import annotations

import sys

import types

import torch

from .. import _lib

P, I32, I64, F64, SZ = _lib.P, _lib.I32, _lib.I64, _lib.F64, _lib.SZ

_lib.register({
    "fc_thermal_rhs": (I32, [P, P, P, P, P, P, P, F64, I64, P]),
    "fc_thermal_reconstruct": (I32, [P, P, P, I32, I32, I32, I32, I32, P]),
    "fc_thermal_metrics_workspace_bytes": (SZ, [I64]),
    "fc_thermal_convergence_metrics": (I32, [P, P, P, I64, P, P, P, SZ, P]),
    "fc_thermal_check_workspace_bytes": (SZ, [I64]),
    "fc_thermal_physics_check": (I32, [P, F64, F64, I64, P, P, SZ, P]),
}, unchecked={"fc_thermal_metrics_workspace_bytes", "fc_thermal_check_workspace_bytes"})


def _p(t, name, dtype=torch.float32, n=None):
    if not t.is_cuda or not t.is_contiguous() or t.dtype != dtype:
        raise RuntimeError(f"{name} must be a contiguous CUDA {dtype} tensor")
    if n is not None and t.numel() != n:
        raise RuntimeError(f"{name} must have {n} elements")
    return t.data_ptr()


def _s(t):
    return torch.cuda.current_stream(t.device).cuda_stream


class _ThermalKernels:
    def thermal_rhs(self, T_old, T0, src_step, LT0, mask_float, rhs, theta0, dt):
        n = T_old.numel()
        _lib.call("fc_thermal_rhs", _p(T_old, "T_old"), _p(T0, "T0", n=n), _p(src_step, "src_step", n=n),
                  _p(LT0, "LT0", n=n), _p(mask_float, "mask_float", n=n), _p(rhs, "rhs", n=n),
                  _p(theta0, "theta0", n=n), float(dt), n, _s(T_old))

    def thermal_reconstruct(self, T0, theta, nx, ny, nz, ax, out_idx):
        T_new = torch.empty_like(T0)
        n = int(nx) * int(ny) * int(nz)
        _lib.call("fc_thermal_reconstruct", _p(T0, "T0", n=n), _p(theta, "theta", n=n), T_new.data_ptr(), int(nx),
                  int(ny), int(nz), int(ax), int(out_idx), _s(T0))
        return T_new

    def convergence_metrics(self, T_new, T_old, solid_u8):
        n = T_new.numel()
        rs = torch.zeros(1, device=T_new.device)
        rf = torch.zeros(1, device=T_new.device)
        ws = torch.empty(int(_lib.call("fc_thermal_metrics_workspace_bytes", n)), dtype=torch.uint8,
                         device=T_new.device)
        _lib.call("fc_thermal_convergence_metrics", _p(T_new, "T_new"), _p(T_old, "T_old", n=n),
                  _p(solid_u8, "solid_u8", torch.uint8, n), n, rs.data_ptr(), rf.data_ptr(), ws.data_ptr(),
                  ws.numel(), _s(T_new))
        return [rs, rf]

    def physics_check(self, T, T_min, T_max):
        T = T.contiguous()
        n = T.numel()
        flag = torch.zeros(1, dtype=torch.int32, device=T.device)
        ws = torch.empty(int(_lib.call("fc_thermal_check_workspace_bytes", n)), dtype=torch.uint8, device=T.device)
        _lib.call("fc_thermal_physics_check", _p(T, "T"), float(T_min), float(T_max), n, flag.data_ptr(),
                  ws.data_ptr(), ws.numel(), _s(T))
        return flag


# As in the reference (thermal_cuda_ext.py:14-15,44-46), a library that cannot be loaded leaves
# THERMAL_CUDA None and THERMAL_CUDA_AVAILABLE False instead of failing the import.
_err = _lib.load_error()
if _err is not None:
    print(f"[THERMAL_CUDA] libfluxcore_hip.so not loaded: {_err}", file=sys.stderr)
    THERMAL_CUDA = None
    THERMAL_CUDA_AVAILABLE = False
else:
    THERMAL_CUDA = _ThermalKernels()
    THERMAL_CUDA_AVAILABLE = bool(torch.cuda.is_available())


def _temperature_step_mg_fused(self, dt: float, tol: float = 1e-3, max_cycles: int = 20) -> dict:
    """One implicit pseudo-time step A theta = rhs, A = I + dt L, theta = T - T0
    (thermal_cuda_ext.py:55-201 / LBM_CHT.py:2571-2712)."""
    from .nlevel_mg_cuda import NLevelGeometricMGSolver

    dt = float(dt)
    T_old = self.T.contiguous()
    if not hasattr(self, "_cached_thermal_setup"):
        update_mask = self._build_update_mask_thermal()
        T0 = self._build_T_dirichlet_field().contiguous()
        src_vol = self._compute_heat_source()
        if src_vol is None:
            src_vol = torch.zeros_like(T_old)
        n_ax = (self.nx, self.ny, self.nz)[self.ax]
        out_idx = n_ax - 1 if self.sign == +1 else 0
        self._cached_thermal_setup = dict(update_mask=update_mask, T0=T0, src_vol=src_vol,
                                          mask_f=update_mask.to(T_old.dtype).contiguous(),
                                          solid_u8=self.solid.to(torch.uint8).contiguous(), out_idx=out_idx)
    cs = self._cached_thermal_setup
    T0, mask_f, out_idx = cs["T0"], cs["mask_f"], cs["out_idx"]
    src_step = (cs["src_vol"] + self._compute_surface_flux_sources()).contiguous()
    if not hasattr(self, "_mg_solver"):
        self._mg_solver = NLevelGeometricMGSolver(
            device=T_old.device, dtype=T_old.dtype, flow_axis=self.ax, flow_sign=self.sign,
            thermal_dt_scale=float(self.thermal_dt_scale), omega=float(getattr(self.cfg, "mg_omega", 0.8)),
            n_pre=int(getattr(self.cfg, "mg_pre_smooth", 2)), n_post=int(getattr(self.cfg, "mg_post_smooth", 2)),
            n_coarse=int(getattr(self.cfg, "mg_coarse_iters", 20)), min_coarse_cells=4)
    self._mg_solver.build_if_needed(k=self.k, rho_cp=self.rho_cp, u=self.u, v=self.v, w=self.wv,
                                    mask=cs["update_mask"], dt=dt)
    # A(T0) = T0 + dt L(T0) in one stencil launch
    LT0 = torch.empty_like(T0)
    self._mg_solver._Mop_cuda(T0, 0, LT0)
    rhs = torch.empty_like(T_old)
    theta0 = torch.empty_like(T_old)
    THERMAL_CUDA.thermal_rhs(T_old, T0, src_step, LT0, mask_f, rhs, theta0, dt)
    theta, mg_info = self._mg_solver.solve(b=rhs, x0=theta0, max_cycles=max_cycles, tol=tol)
    Tin = float(self.fluid.tin_C)
    theta.mul_(mask_f)
    flag = THERMAL_CUDA.physics_check(T0 + theta, Tin - 500.0, Tin + 2000.0)
    T_new = THERMAL_CUDA.thermal_reconstruct(T0, theta, self.nx, self.ny, self.nz, self.ax, out_idx)
    if int(flag.item()):
        return dict(converged=False, it=mg_info["it"], res_rel=mg_info["res_rel"], r_rel=float("inf"),
                    r_inf=float("inf"), src_inf=float("nan"), Tmax=float("nan"), diverged=True)
    self.T = T_new
    rs, rf = THERMAL_CUDA.convergence_metrics(T_new, T_old, cs["solid_u8"])
    self._async_dT_solid, self._async_dT_fluid = rs, rf
    with torch.no_grad():
        LT = self._mg_solver.L_apply_fine(self.T)
        src_ss = src_step / dt
        r_ss = (LT - src_ss) * mask_f
        r_inf = float(r_ss.abs().max())
        m = cs["update_mask"]
        src_inf = float(src_ss[m].abs().max()) if m.any() else 1e-30
    return dict(converged=mg_info["converged"], it=mg_info["it"], res_rel=mg_info["res_rel"],
                r_rel=r_inf / max(src_inf, 1e-30), r_inf=r_inf, src_inf=src_inf, Tmax=float(self.T.max()))


def patch_thermal_fused(sim):
    """Route sim._temperature_step_mg through the fused kernels (thermal_cuda_ext.py:204-234)."""
    sim._temperature_step_mg = types.MethodType(_temperature_step_mg_fused, sim)
    return sim
