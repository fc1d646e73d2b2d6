"""Conduction problems shared by the MG tests and the thermal benchmark."""
import numpy as np

from fluxcore3d_amd import geometry


def sphere_conduction(nx=64, ny=32, nz=32, k_f=0.02, ratio=100.0, dt=1.0e6):
    """Config-1 conduction: k_s/k_f = ratio in the r=6 sphere, uniform rho_cp,
    u = 0, T = 0 (Dirichlet, mask 0) on the +-Y/+-Z wall planes, unit total
    source spread over the sphere cells, implicit step dt (steady limit)."""
    phi, solid, _ = geometry.sphere_channel(nx, ny, nz, center=(20.0 * nx / 64, ny / 2, nz / 2), radius=6.0 * ny / 32)
    k = np.where(solid, k_f * ratio, k_f).astype(np.float32)
    rcp = np.ones_like(k)
    z = np.zeros_like(k)
    mask = np.ones_like(k, dtype=bool)
    mask[:, 0, :] = mask[:, -1, :] = False
    mask[:, :, 0] = mask[:, :, -1] = False
    src = np.where(solid, 1.0 / max(int(solid.sum()), 1), 0.0).astype(np.float32)
    b = (dt * src).astype(np.float32) * mask
    return dict(k=k, rcp=rcp, u=z, v=z, w=z, mask=mask, b=b, dt=dt, ax=0, sign=1, s=1.0)


def _host(T):
    """Host FP64 copy of a field held as a torch tensor or a NumPy array."""
    return T.double().cpu().numpy() if hasattr(T, "cpu") else np.asarray(T, np.float64)


def steady_thermal_loop(sim, solid, max_outer=500, tol_dT_solid=1e-2, tol_dT_fluid=1e-2, tol_dTout_mean=5e-3,
                        dt_scale_min=1.0, dt_scale_max=200.0, dt_scale_start=5.0, dt_ramp_factor=1.5,
                        dt_backoff_factor=0.5, dt_ramp_threshold=0.5, max_mg_cycles=20, tol_mg=1e-3,
                        stable_steps=5, outlet=None, before_step=None):
    """The pseudo-time outer loop of solve_thermal_steady_only (LBM_CHT.py:2388-2569)
    driving ``sim._temperature_step_mg`` (the build's fused driver, or the oracle's
    restatement) on host copies of the fields:
      * a step reported as diverged (the physics check) leaves T unchanged, zeroes the
        multigrid iterate (when the simulation has one) and halves dt (not below
        dt_scale_min);
      * a non-finite T or max T above 1e6 restores the previous T and halves dt;
      * dt grows by dt_ramp_factor (up to dt_scale_max) when the step took fewer than
        dt_ramp_threshold * max_mg_cycles cycles;
      * after an accepted step the simulation's _apply_temperature_bcs runs (the face
        conditions of LBM_CHT.py:1946-2015, see apply_temperature_bcs);
      * the loop stops after ``stable_steps`` consecutive steps with max |dT| over solid
        and fluid and the change of the mean outlet temperature below their tolerances.
    ``outlet`` = (axis, index) of the outlet plane; ``before_step(outer, sim)`` lets a
    test change the simulation before a step. Returns the per-step log and a summary."""
    fluid = ~solid
    dt_use = float(dt_scale_start)
    Tout_prev = None
    stable = 0
    log = []
    for outer in range(max_outer):
        if before_step is not None:
            before_step(outer, sim)
        Tprev = sim.T.clone() if hasattr(sim.T, "clone") else sim.T.copy()
        info = sim._temperature_step_mg(dt=float(dt_use), tol=tol_mg, max_cycles=max_mg_cycles)
        mg_it = info.get("it", max_mg_cycles)
        if info.get("diverged", False):
            if hasattr(sim, "_mg_solver") and sim._mg_solver._ready:
                sim._mg_solver._levels[0]["x"].zero_()
            log.append(dict(outer=outer, dt=dt_use, status="backoff", it=mg_it))
            dt_use = max(float(dt_scale_min), dt_backoff_factor * dt_use)
            continue
        Th = _host(sim.T)
        Tmax_new = float(Th.max())
        if not np.isfinite(Th).all() or Tmax_new > 1e6:
            sim.T = Tprev
            log.append(dict(outer=outer, dt=dt_use, status="restored", it=mg_it))
            dt_use = max(float(dt_scale_min), dt_backoff_factor * dt_use)
            continue
        log.append(dict(outer=outer, dt=dt_use, status="accepted", it=mg_it))
        if mg_it < dt_ramp_threshold * max_mg_cycles:
            dt_use = min(float(dt_scale_max), dt_ramp_factor * dt_use)
        sim._apply_temperature_bcs()
        Th = _host(sim.T)
        dT = np.abs(Th - _host(Tprev))
        dT_s = float(dT[solid].max()) if solid.any() else 0.0
        dT_f = float(dT[fluid].max()) if fluid.any() else 0.0
        if outlet is not None:
            sl = [slice(None)] * 3
            sl[outlet[0]] = outlet[1]
            fs = fluid[tuple(sl)]
            Tout = float(Th[tuple(sl)][fs].mean()) if fs.any() else float(Th[fluid].mean())
        else:
            Tout = float(Th[fluid].mean())
        dTout = float("inf") if Tout_prev is None else abs(Tout - Tout_prev)
        Tout_prev = Tout
        log[-1].update(dT_s=dT_s, dT_f=dT_f, dTout=dTout)
        if dT_s < tol_dT_solid and dT_f < tol_dT_fluid and dTout < tol_dTout_mean:
            stable += 1
            if stable >= stable_steps:
                return log, dict(outer_iters=outer + 1, Tmax=Tmax_new, Tout_mean=Tout, converged=True)
    return log, dict(outer_iters=max_outer, Tmax=float(_host(sim.T).max()), converged=False)


def apply_temperature_bcs(T, axis, sign, Tin):
    """_apply_temperature_bcs (LBM_CHT.py:1946-1993, 1995-2015), in place on a torch
    tensor or a NumPy array: transverse domain faces copy their inner neighbour (x, then
    y, then z), the inlet plane is set to Tin, the outlet plane copies its inner
    neighbour."""
    n = T.shape[axis]
    inlet_idx, outlet_idx = (0, n - 1) if sign == +1 else (n - 1, 0)
    for a in range(3):
        if a == axis:
            continue
        lo, lo1, hi, hi1 = ([slice(None)] * 3 for _ in range(4))
        lo[a], lo1[a], hi[a], hi1[a] = 0, 1, -1, -2
        T[tuple(lo)] = T[tuple(lo1)]
        T[tuple(hi)] = T[tuple(hi1)]
    sl = [slice(None)] * 3
    sl[axis] = inlet_idx
    T[tuple(sl)] = Tin
    nei = outlet_idx - 1 if outlet_idx > 0 else outlet_idx + 1
    so, sn = [slice(None)] * 3, [slice(None)] * 3
    so[axis], sn[axis] = outlet_idx, nei
    T[tuple(so)] = T[tuple(sn)]


class OracleThermalSim:
    """The oracle's side of steady_thermal_loop: the FP32 NumPy hierarchy and
    oracle.mg.thermal_step_guarded with the fused driver's bounds [Tin - 500, Tin + 2000]
    (thermal_cuda_ext.py:146-160)."""

    def __init__(self, h, T, T0, src, mask, tin_C):
        from types import SimpleNamespace
        self.h, self.T, self.T0, self.src, self.mask = h, T, T0, src, mask
        self.fluid = SimpleNamespace(tin_C=tin_C)

    def _apply_temperature_bcs(self):
        apply_temperature_bcs(self.T, self.h.ax, self.h.sign, self.fluid.tin_C)

    def _temperature_step_mg(self, dt, tol, max_cycles):
        from oracle import mg
        self.h.set_dt(dt)
        T_new, hist = mg.thermal_step_guarded(self.h, self.T, self.T0, self.src, self.mask, dt,
                                              self.fluid.tin_C - 500.0, self.fluid.tin_C + 2000.0, tol=tol,
                                              max_cycles=max_cycles)
        if T_new is None:
            return dict(converged=False, it=len(hist), diverged=True)
        self.T = T_new
        return dict(converged=hist[-1] < tol, it=len(hist))
