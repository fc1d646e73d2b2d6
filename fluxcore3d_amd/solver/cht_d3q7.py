"""Coupled conjugate heat transfer on the flow's lattice: D3Q7 fluid temperature
every flow step, implicit conduction of the solid every ``couple_every`` steps.

No counterpart in the reference. Its pipeline (UI_components/cht_worker.py:23-165)
solves the flow to steady state, freezes it and then marches the temperature of
fluid and solid together with the implicit finite-volume upwind operator
(solver/nlevel_mg_cuda.py:418-507, LBM_CHT.py:2208-2359, driven by
solve_thermal_steady_only, LBM_CHT.py:2485-2558). Here the fluid temperature is
carried by the flow as it develops:

* ``D3Q7Transport`` — one HIP launch per flow step (fc_d3q7_step): BGK D3Q7
  advection-diffusion with the step's velocity, c_s^2 = 1/4, kappa_f = (1/omega -
  1/2)/4; adiabatic walls bounce back, conducting solids enter through
  anti-bounce-back at their temperature, inlet nodes hold the inlet temperature,
  open faces have zero gradient. In the same launch every conducting node adds
  the energy its fluid neighbours hand to it.
* ``SolidConduction`` — every ``couple_every`` flow steps one implicit step of the
  solid over that interval: rho_cp_s (T - T_old) + dt K T = acc + dt rho_cp_s src,
  K the face Laplacian of conductance k_s between conducting cells (zero flux to
  everything else), acc the accumulated interface energy, src the per-step
  temperature source (a surface flux as in SurfaceFluxBC.src_step,
  thermal_bc.py:170-196). Solved with flexible CG in FP64 preconditioned by the
  FP32 geometric-multigrid V-cycle (A = I + dt L of NLevelGeometricMGSolver with
  k = k_s on conducting cells and 0 elsewhere, so the harmonic face mean
  decouples the solid), on a box one cell larger than the solid on every side so
  that no conducting cell touches the multigrid's mirrored domain faces.
* ``CoupledCHT`` — the loop around a drop-in flow simulation
  (LBMCUDAUpgrade.patch): flow step -> D3Q7 step -> solid step every
  ``couple_every`` steps, with the callbacks of run_simulation_blocking.

The exchange is conservative by construction: the solid receives exactly the
energy the fluid's anti-bounce-back inputs remove (same operands, same order),
so fluid energy + accumulated energy changes only through the inlet, the outlet
and the source. Temperatures are carried as deviations from the inlet
temperature; ``temperature()`` adds it back.
"""
from __future__ import annotations

import threading
from collections import deque
from typing import Callable, Optional

import torch

from .. import _lib
from . import mg_cuda_ext  # noqa: F401  (registers the multigrid entry points)
from .mg_pcg import FCGWorkspace, fcg_fp64
from .nlevel_mg_cuda import NLevelGeometricMGSolver

Tensor = torch.Tensor
P, I32, I64, F32, F64 = _lib.P, _lib.I32, _lib.I64, _lib.F32, _lib.F64

_lib.register({
    "fc_d3q7_step": (I32, [P, P, P, P, P, P, P, P, P, P, P, I32, I32, I32, I32, I32, I32, F32, F32, F32, P]),
    "fc_d3q7_init": (I32, [P, P, P, P, P, I64, P]),
    "fc_cht_solid_rhs": (I32, [P, P, P, P, P, P, I32, I32, I32, I32, I32, I32, I32, I32, I32, F64, F64, P, P, P]),
    "fc_cht_solid_store": (I32, [P, P, P, P, I32, I32, I32, I32, I32, I32, I32, I32, I32, P]),
})

FLUID, SOLID, WALL, CONJ, INLET = 0, 1, 2, 3, 16


def _shifted(field: Tensor, ax: int, c: int, periodic) -> Tensor:
    """Value at x of the neighbour x - c e_ax (0 where that leaves a non-periodic axis)."""
    nb = torch.roll(field, shifts=c, dims=ax)
    if not periodic[ax]:
        idx = [slice(None)] * 3
        idx[ax] = 0 if c > 0 else -1
        nb[tuple(idx)] = 0
    return nb


def neighbour_mask(flag: Tensor, periodic) -> Tensor:
    """uint8 per node, bits q - 1 for q = 1..6 (+x, -x, +y, -y, +z, -z); neighbours inside
    the domain, wrapped on periodic axes.
    * Solid nodes solved apart (kind 1): x + c_q is a fluid node off the inlet (the
      neighbours whose energy the solid node receives).
    * Every other node: x - c_q is a wall or a solid node solved apart (the inputs the
      D3Q7 step replaces; it reads only those neighbours' flags)."""
    kind = flag.to(torch.int32) & 3
    sw = ((kind == SOLID) | (kind == WALL)).to(torch.int32)
    fl_open = ((kind == FLUID) & ((flag.to(torch.int32) & INLET) == 0)).to(torch.int32)
    pull = torch.zeros_like(sw)
    give = torch.zeros_like(sw)
    for q, (ax, c) in enumerate(((0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1))):
        pull |= _shifted(sw, ax, c, periodic) << q         # x - c_q
        give |= _shifted(fl_open, ax, -c, periodic) << q   # x + c_q
    return torch.where(kind == SOLID, give, pull).to(torch.uint8).contiguous()


def _stream(t: Tensor):
    return torch.cuda.current_stream(t.device).cuda_stream


def _ptr(t: Optional[Tensor]):
    return None if t is None else t.data_ptr()


def _f32(t: Tensor, n: int, name: str) -> Tensor:
    if t.dtype != torch.float32 or not t.is_contiguous() or t.numel() != n or t.device.type != "cuda":
        raise ValueError(f"{name}: expected a contiguous float32 HIP tensor of {n} elements")
    return t


def node_flags(conducting: Tensor, wall: Optional[Tensor] = None, inlet: Optional[Tensor] = None) -> Tensor:
    """uint8 node kinds: 1 conducting solid, 2 adiabatic wall (non-conducting solid of
    the flow), 0 fluid, | 16 on fluid nodes of the inlet plane."""
    fl = torch.zeros(conducting.shape, dtype=torch.uint8, device=conducting.device)
    if wall is not None:
        fl[wall.bool()] = WALL
    fl[conducting.bool()] = SOLID
    if inlet is not None:
        fl[inlet.bool() & (fl == FLUID)] = INLET
    return fl.contiguous()


class D3Q7Transport:
    """Fluid temperature populations (two [7][N] FP32 buffers) and the step launch.
    ``omega_s`` relaxes kind-3 nodes (conducting solid carried on the lattice). The node
    flags are fixed for the transport's life: the neighbour mask is built from them once."""

    def __init__(self, flag: Tensor, omega: float, periodic=(0, 0, 0), T_in: float = 0.0,
                 omega_s: Optional[float] = None):
        if flag.dtype != torch.uint8 or flag.dim() != 3 or flag.device.type != "cuda":
            raise ValueError("flag: expected a 3-D uint8 HIP tensor")
        omega_s = omega if omega_s is None else omega_s
        if not (0.0 < float(omega) < 2.0 and 0.0 < float(omega_s) < 2.0):
            raise ValueError("omega and omega_s must be in (0, 2)")
        err = _lib.load_error()
        if err is not None:
            raise RuntimeError(f"D3Q7Transport needs libfluxcore_hip.so: {err}")
        self.flag = flag.contiguous()
        self.shape = tuple(flag.shape)
        self.N = int(flag.numel())
        self.omega = float(omega)
        self.omega_s = float(omega_s)
        self.periodic = tuple(int(bool(p)) for p in periodic)
        self.T_in = float(T_in)
        dev = flag.device
        self.nbmask: Optional[Tensor] = neighbour_mask(self.flag, self.periodic)  # None: every flag read
        self.g = [torch.zeros((7, self.N), dtype=torch.float32, device=dev) for _ in range(2)]
        self.T = torch.zeros(self.N, dtype=torch.float32, device=dev)
        self.cur = 0

    @property
    def kappa(self) -> float:
        return (1.0 / self.omega - 0.5) / 4.0

    def init(self, T0: Tensor, u: Tensor, v: Tensor, w: Tensor) -> None:
        """Populations at equilibrium with temperature T0 and velocity (u, v, w)."""
        args = [_f32(t.reshape(-1), self.N, n) for t, n in ((T0, "T0"), (u, "u"), (v, "v"), (w, "w"))]
        self.T.copy_(args[0])
        _lib.call("fc_d3q7_init", self.g[self.cur].data_ptr(), *[a.data_ptr() for a in args], self.N,
                  _stream(self.T))

    def step(self, u: Tensor, v: Tensor, w: Tensor, T_solid: Tensor, acc: Optional[Tensor] = None,
             Q: Optional[Tensor] = None) -> Tensor:
        """One step with this flow step's velocity; returns T (fluid from D3Q7,
        conducting nodes T_solid)."""
        uu = [_f32(t.reshape(-1), self.N, n) for t, n in ((u, "u"), (v, "v"), (w, "w"))]
        ts = _f32(T_solid.reshape(-1), self.N, "T_solid")
        a = None if acc is None else _f32(acc.reshape(-1), self.N, "acc")
        q = None if Q is None else _f32(Q.reshape(-1), self.N, "Q")
        nx, ny, nz = self.shape
        src, dst = self.g[self.cur], self.g[1 - self.cur]
        _lib.call("fc_d3q7_step", src.data_ptr(), dst.data_ptr(), self.T.data_ptr(), *[t.data_ptr() for t in uu],
                  self.flag.data_ptr(), _ptr(self.nbmask), ts.data_ptr(), _ptr(q), _ptr(a), nx, ny, nz,
                  *self.periodic, self.omega,
                  self.omega_s, self.T_in, _stream(self.T))
        self.cur = 1 - self.cur
        return self.T

    @property
    def populations(self) -> Tensor:
        """Post-collision populations of the last step ([7][N])."""
        return self.g[self.cur]


class SolidConduction:
    """Implicit conduction step of the conducting nodes over ``dt`` flow steps."""

    def __init__(self, flag: Tensor, k_s: float, rho_cp: float, dt: float, tol: float = 1e-8, max_it: int = 300,
                 omega: float = 0.8, n_pre: int = 3, n_post: int = 3, skip_empty_rows: bool = True):
        if k_s <= 0 or rho_cp <= 0 or dt <= 0:
            raise ValueError("k_s, rho_cp and dt must be positive")
        self.flag = flag.contiguous()
        self.shape = tuple(flag.shape)
        self.k_s, self.rho_cp, self.dt, self.tol, self.max_it = float(k_s), float(rho_cp), float(dt), tol, max_it
        dev = flag.device
        cond = (flag & 3) == SOLID
        if not bool(cond.any()):
            raise ValueError("no conducting nodes")
        lo, hi = [], []
        for ax in range(3):
            other = tuple(a for a in range(3) if a != ax)
            idx = torch.nonzero(cond.any(dim=other[1]).any(dim=other[0])).reshape(-1)
            lo.append(int(idx.min()))
            hi.append(int(idx.max()))
        # one non-conducting cell on every side, extents rounded up to multiples of 8
        self.origin = tuple(v - 1 for v in lo)
        self.box = tuple(((h - l + 3) + 7) // 8 * 8 for l, h in zip(lo, hi))
        box_cond = torch.zeros(self.box, dtype=torch.bool, device=dev)
        g_sl, b_sl = [], []
        for ax in range(3):
            g0 = max(self.origin[ax], 0)
            g1 = min(self.origin[ax] + self.box[ax], self.shape[ax])
            g_sl.append(slice(g0, g1))
            b_sl.append(slice(g0 - self.origin[ax], g1 - self.origin[ax]))
        box_cond[tuple(b_sl)] = cond[tuple(g_sl)]
        self.n_unknowns = int(box_cond.sum())
        # conductance k_s inside the solid, 0 outside (the harmonic face mean then decouples
        # the solid); heat capacity rho_cp everywhere, so partially solid coarse cells
        # keep the solid's; coarse unknowns wherever any child is solid (thin pins stay
        # on the coarse levels: 55 instead of 233 FCG iterations on the quarter-scale
        # heat sink in the FP64 model of the same hierarchy)
        k = torch.where(box_cond, torch.tensor(self.k_s, device=dev), torch.tensor(0.0, device=dev))
        rcp = torch.full(self.box, self.rho_cp, device=dev)
        z = torch.zeros(self.box, dtype=torch.float32, device=dev)
        self.mg = NLevelGeometricMGSolver(dev, torch.float32, flow_axis=2, flow_sign=+1, thermal_dt_scale=1.0,
                                          omega=omega, n_pre=n_pre, n_post=n_post, coarse_mask="any",
                                          skip_empty_rows=skip_empty_rows)
        self.mg.build_if_needed(k.float(), rcp.float(), z, z, z, box_cond.float(), self.dt)
        self.b = torch.zeros(self.box, dtype=torch.float64, device=dev)
        self.x0 = torch.zeros(self.box, dtype=torch.float64, device=dev)
        self.ws = FCGWorkspace(self.box, dev)
        # start from 2 T_s - T_prev after the first step and from 3 (T_s - T_prev) + T_prev2
        # after the second (T_prev, T_prev2: the previous intervals' start temperatures);
        # on the full-size heat sink over 10 000 steps 27.0 FCG iterations per solve on
        # average instead of 36.5 with the linear start
        self.extrapolate = True
        self._prev: Optional[Tensor] = None
        self._prev2: Optional[Tensor] = None

    def _geom(self):
        return (*self.shape, *self.origin, *self.box)

    def step(self, T_solid: Tensor, acc: Tensor, src: Optional[Tensor] = None) -> dict:
        """T_solid <- solution of the implicit step; acc restarts at zero."""
        n = 1
        for s in self.shape:
            n *= s
        ts = _f32(T_solid.reshape(-1), n, "T_solid")
        a = _f32(acc.reshape(-1), n, "acc")
        q = None if src is None else _f32(src.reshape(-1), n, "src")
        stream = _stream(ts)
        prev = self._prev if self.extrapolate else None
        prev2 = self._prev2 if self.extrapolate else None
        _lib.call("fc_cht_solid_rhs", ts.data_ptr(), _ptr(prev), _ptr(prev2), a.data_ptr(), _ptr(q),
                  self.flag.data_ptr(),
                  *self._geom(), self.rho_cp, self.dt, self.b.data_ptr(), self.x0.data_ptr(), stream)
        _, info = fcg_fp64(self.mg, self.b, x0=self.x0, tol=self.tol, max_it=self.max_it, ws=self.ws)
        if not info["converged"]:
            raise RuntimeError(f"solid conduction step did not reach {self.tol:g} in {info['it']} iterations "
                               f"(relative residual {info['res_rel']:.3e})")
        if self.extrapolate:  # this interval's start temperature, for the next start guesses
            if self._prev is None:
                self._prev = torch.empty_like(ts)
            elif self._prev2 is None:
                self._prev2 = self._prev.clone()
            else:
                self._prev2.copy_(self._prev)
            self._prev.copy_(ts)
        _lib.call("fc_cht_solid_store", info["x64"].data_ptr(), ts.data_ptr(), a.data_ptr(), self.flag.data_ptr(),
                  *self._geom(), stream)
        return {"it": info["it"], "res_rel": info["res_rel"]}


def inlet_mask(sim) -> Optional[Tensor]:
    """Fluid nodes of the velocity-inlet plane of the drop-in simulation's flow plan."""
    from .lbm_cuda_ext import flow_bc_plan
    m = None
    for bc in flow_bc_plan(sim):
        if bc["kind"] == "zou_he" and bc["bc_type"] == 0:
            m = torch.zeros((sim.nx, sim.ny, sim.nz), dtype=torch.bool, device=sim.solid.device) if m is None else m
            sl = [slice(None)] * 3
            sl[bc["axis"]] = bc["plane"]
            m[tuple(sl)] = True
    return m


class CoupledCHT:
    """Flow step -> D3Q7 step -> (every ``couple_every`` steps) solid step.

    ``sim``: a drop-in flow simulation (LBMCHT3D_Torch fields, LBMCUDAUpgrade.patch
    applied) whose ``step(do_thermal=False)`` leaves the step's velocity in
    ``sim.u``, ``sim.v``, ``sim.wv``. Conducting nodes are ``sim.solid``, adiabatic
    walls the rest of ``sim.solid_flow``. Fluid rho c_p is the unit; the solid has
    conductivity ``k_ratio`` kappa_f and heat capacity ``rho_cp_ratio``; kappa_f =
    nu / ``prandtl`` with nu from the flow's relaxation rate. ``src``: per-step
    temperature source of the solid (K per flow step, solid heat capacity).

    ``overlap`` (fused AA engine with its plain step only): the temperature does not act on the flow, so the
    D3Q7 and solid steps run on a second stream, and the solid solve (small kernels, the
    host waiting on each iteration's residual) runs from a host thread while the flow steps
    go on. The velocity goes through a ring of ``ring`` buffer sets (``sim.u``, ``sim.v``,
    ``sim.wv`` the first); a flow step waits only for the D3Q7 step that read its set
    ``ring`` steps earlier, and D3Q7 steps due during a solve are queued behind it. Same
    kernels and inputs in the same order on each stream, so the same bits as the
    sequential loop; ``sync()`` (called by ``temperature()`` and ``run()``) joins the
    thread and the streams and leaves the last step's velocity in the ``sim`` fields."""

    def __init__(self, sim, src: Optional[Tensor] = None, k_ratio: float = 1000.0, rho_cp_ratio: float = 1.6,
                 prandtl: float = 0.71, couple_every: int = 50, tol: float = 1e-8, T_in: float = 0.0,
                 periodic=(0, 0, 0), overlap: bool = True, ring: int = 48):
        self.sim = sim
        self.every = int(couple_every)
        if self.every < 1:
            raise ValueError("couple_every must be >= 1")
        nu = (1.0 / float(sim.omega_plus) - 0.5) / 3.0
        self.kappa_f = nu / float(prandtl)
        self.omega_g = 1.0 / (0.5 + 4.0 * self.kappa_f)
        self.k_s = float(k_ratio) * self.kappa_f
        self.rho_cp_s = float(rho_cp_ratio)
        self.T_in = float(T_in)
        wall = sim.solid_flow.bool() & ~sim.solid.bool()
        self.flag = node_flags(sim.solid, wall, inlet_mask(sim))
        self.transport = D3Q7Transport(self.flag, self.omega_g, periodic=periodic, T_in=0.0)
        self.solid = SolidConduction(self.flag, self.k_s, self.rho_cp_s, float(self.every), tol=tol)
        N = self.transport.N
        dev = self.flag.device
        self.T_solid = torch.zeros(N, dtype=torch.float32, device=dev)
        self.acc = torch.zeros(N, dtype=torch.float32, device=dev)
        self.src = None if src is None else src.to(device=dev, dtype=torch.float32).reshape(-1).contiguous()
        z = torch.zeros(N, dtype=torch.float32, device=dev)
        self.transport.init(z, z, z, z)
        self.n = 0
        self.history: list = []
        self._eng = sim.__dict__.get("_fc_engine")
        from .lbm_cuda_ext import LBMCUDAUpgrade
        plain = getattr(sim.step, "__func__", None) is LBMCUDAUpgrade._step_aa  # not a profiling wrapper
        self.overlap = (bool(overlap) and self._eng is not None and plain
                        and not getattr(sim, "_cuda_debug_nan", False))
        if self.overlap:
            if int(ring) < 2:
                raise ValueError("ring must be >= 2")
            self._ring = int(ring)
            self._tstream = torch.cuda.Stream(device=dev)  # default priority, like every stream here
            self._vel = [tuple(t.view(-1) for t in (sim.u, sim.v, sim.wv))]
            self._vel += [tuple(torch.zeros(N, dtype=torch.float32, device=dev) for _ in range(3))
                          for _ in range(self._ring - 1)]
            self._set_ready = [False] * self._ring  # solid entries written into the set
            self._read_done: list = [None] * self._ring  # event: the thermal stream is done with the set
            self._flow_done: list = [None] * self._ring  # event: the flow step that wrote the set
            self._pending: deque = deque()  # flow steps whose D3Q7 step is not queued yet
            self._n_thermal = 0
            self._solver: Optional[threading.Thread] = None
            self._solver_error: Optional[BaseException] = None
            self._last = -1

    def step(self) -> None:
        if self.overlap:
            self._step_overlapped()
            return
        sim = self.sim
        sim.step(do_thermal=False)
        self.transport.step(sim.u, sim.v, sim.wv, self.T_solid, acc=self.acc)
        self.n += 1
        if self.n % self.every == 0:
            info = self.solid.step(self.T_solid, self.acc, self.src)
            info["step"] = self.n
            self.history.append(info)

    def _step_overlapped(self) -> None:
        k = self.n % self._ring
        while self._pending and self._pending[0] <= self.n - self._ring:
            self._join_solver()  # set k is still to be read by a D3Q7 step not queued yet
            self._flush()
        flow = torch.cuda.current_stream(self.T_solid.device)
        if self._read_done[k] is not None:  # the D3Q7 step that last read set k is done with it
            flow.wait_event(self._read_done[k])
        u, v, w = self._vel[k]
        self._eng.step(self.sim.rho.view(-1), u, v, w, solid_macros=not self._set_ready[k])
        self._set_ready[k] = True
        done = torch.cuda.Event()
        done.record(flow)
        self._flow_done[k] = done
        self._pending.append(self.n)
        self._last = k
        self.n += 1
        self._flush()

    def _flush(self) -> None:
        """Queue the pending D3Q7 steps on the thermal stream, in order, unless a solid
        solve is being queued there; a step that completes a coupling interval starts its
        solve in a host thread."""
        if self._solver is not None and self._solver.is_alive():
            return
        self._join_solver()
        while self._pending:
            m = self._pending.popleft()
            k = m % self._ring
            self._tstream.wait_event(self._flow_done[k])
            u, v, w = self._vel[k]
            with torch.cuda.stream(self._tstream):
                self.transport.step(u, v, w, self.T_solid, acc=self.acc)
                done = torch.cuda.Event()
                done.record(self._tstream)
            self._read_done[k] = done
            self._n_thermal += 1
            if self._n_thermal % self.every == 0:
                self._solver = threading.Thread(target=self._solve, args=(self._n_thermal,), daemon=True)
                self._solver.start()
                return

    def _solve(self, step: int) -> None:
        try:
            with torch.cuda.stream(self._tstream):
                info = self.solid.step(self.T_solid, self.acc, self.src)
            info["step"] = step
            self.history.append(info)
        except BaseException as e:  # re-raised in the calling thread by _join_solver
            self._solver_error = e

    def _join_solver(self) -> None:
        if self._solver is not None:
            self._solver.join()
            self._solver = None
        if self._solver_error is not None:
            e, self._solver_error = self._solver_error, None
            raise e

    def sync(self) -> None:
        """Finish queuing every step, join the thermal stream into the current one, and
        leave the last step's velocity in ``sim.u``, ``sim.v``, ``sim.wv`` (no-op for the
        sequential loop)."""
        if not self.overlap or self._last < 0:
            return
        while True:
            self._join_solver()
            self._flush()
            if not self._pending and self._solver is None:
                break
        self._join_solver()
        torch.cuda.current_stream(self.T_solid.device).wait_stream(self._tstream)
        if self._last != 0:
            for dst, src in zip(self._vel[0], self._vel[self._last]):
                dst.copy_(src)

    def run(self, steps: int, log_fn: Optional[Callable] = None, progress_fn: Optional[Callable] = None,
            abort_fn: Optional[Callable] = None, report_every: int = 1000) -> dict:
        """``steps`` coupled steps with the callbacks of run_simulation_blocking
        (log_fn(str), progress_fn(0..100), abort_fn() -> True to stop)."""
        for s in range(int(steps)):
            if abort_fn is not None and abort_fn():
                self.sync()  # the steps taken so far complete, the solve thread is joined
                raise RuntimeError("Aborted by user.")
            self.step()
            if progress_fn is not None and (s + 1) % max(1, steps // 100) == 0:
                progress_fn(int(100 * (s + 1) / steps))
            if log_fn is not None and (s + 1) % report_every == 0:
                self.sync()
                last = self.history[-1] if self.history else {}
                log_fn(f"[CHT] step {self.n}: T_max {float(self.temperature().max()):.4f}, "
                       f"solid solve {last.get('it', 0)} it, residual {last.get('res_rel', 0.0):.2e}")
        self.sync()
        return {"steps": self.n, "solves": len(self.history)}

    def temperature(self) -> Tensor:
        """(nx, ny, nz) temperature: fluid from D3Q7, conducting nodes from the solid."""
        self.sync()
        T = torch.where((self.flag.reshape(-1) & 3) == SOLID, self.T_solid, self.transport.T)
        return (T + self.T_in).view(self.transport.shape)


# ----------------------------------------------------------------------------
# Conjugate heat on periodic porous media: the solid carried on the D3Q7 lattice
# (kind 3, u = 0, its own relaxation rate), one launch per step for fluid and solid.
# ----------------------------------------------------------------------------
_C7 = ((0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))
_OPP7 = (0, 2, 1, 4, 3, 6, 5)


def _interface_masks(flag: Tensor, solid_kind: int):
    kind = flag & 3
    fluid = kind == FLUID
    solid = kind == solid_kind
    return [fluid & torch.roll(solid, shifts=_C7[q], dims=(0, 1, 2)) for q in range(7)]


def interface_heat(g: Tensor, flag: Tensor, solid_kind: int = CONJ) -> float:
    """Energy per step crossing from solid nodes into fluid nodes of a periodic
    domain, from the post-collision populations g ([7][N]): over every fluid node x
    and direction q with x - c_q solid, g*_q(x - c_q) - g*_opp(q)(x) (FP64 sums)."""
    shape = tuple(flag.shape)
    gv = g.view(7, *shape)
    masks = _interface_masks(flag, solid_kind)
    e = torch.zeros((), dtype=torch.float64, device=g.device)
    for q in range(1, 7):
        m = masks[q]
        e += torch.roll(gv[q], shifts=_C7[q], dims=(0, 1, 2))[m].double().sum() - gv[_OPP7[q]][m].double().sum()
    return float(e)


def interface_links(flag: Tensor, solid_kind: int = CONJ) -> int:
    """Fluid-solid faces of a periodic domain (links between face neighbours)."""
    return int(sum(int(m.sum()) for m in _interface_masks(flag, solid_kind)[1:]))


def nusselt(g: Tensor, T: Tensor, flag: Tensor, kappa_f: float, solid_kind: int = CONJ) -> dict:
    """Interstitial Nusselt number of a periodic conjugate problem: h = Q / (A (mean
    T_solid - mean T_fluid)), Nu = h d_h / k_f, Q the energy per step crossing the
    interface, A the interface links, d_h = 4 V_fluid / A, k_f = kappa_f (fluid
    rho c_p = 1)."""
    kind = flag.reshape(-1) & 3
    Tf64 = T.reshape(-1).double()
    Q = interface_heat(g, flag, solid_kind)
    A = interface_links(flag, solid_kind)
    Ts = float(Tf64[kind == solid_kind].mean())
    Tfl = float(Tf64[kind == FLUID].mean())
    d_h = 4.0 * float((kind == FLUID).sum()) / A
    h = Q / (A * (Ts - Tfl))
    return {"Nu": h * d_h / kappa_f, "h": h, "Q": Q, "A": A, "d_h": d_h, "T_solid_mean": Ts, "T_fluid_mean": Tfl}


class ConjugateHeat:
    """Periodic flow (an AAEngine, one fused launch per step) and conjugate heat:
    fluid temperature on D3Q7 with the step's velocity, the solid carried on the
    same lattice with conductivity ``k_ratio`` kappa_f (equal rho c_p), a uniform
    source ``q_solid`` per solid node per step and the balancing uniform sink in the
    fluid (a statistically steady state without drift of the mean temperature)."""

    def __init__(self, engine, solid: Tensor, kappa_f: float, k_ratio: float = 10.0, q_solid: float = 1e-4):
        dev = solid.device
        self.engine = engine
        self.shape = tuple(solid.shape)
        N = solid.numel()
        self.flag = torch.where(solid.bool(), torch.tensor(CONJ, dtype=torch.uint8, device=dev),
                                torch.tensor(FLUID, dtype=torch.uint8, device=dev)).contiguous()
        self.kappa_f = float(kappa_f)
        self.kappa_s = float(k_ratio) * self.kappa_f
        omega_f = 1.0 / (0.5 + 4.0 * self.kappa_f)
        omega_s = 1.0 / (0.5 + 4.0 * self.kappa_s)
        self.transport = D3Q7Transport(self.flag, omega_f, periodic=(1, 1, 1), omega_s=omega_s)
        n_s = int(solid.sum())
        n_f = N - n_s
        self.Q = torch.where(solid.reshape(-1).bool(), torch.tensor(float(q_solid), device=dev),
                             torch.tensor(-float(q_solid) * n_s / n_f, device=dev)).contiguous()
        self.macro = [torch.zeros(N, device=dev) for _ in range(4)]
        self._zero = torch.zeros(N, device=dev)
        self.transport.init(self._zero, self._zero, self._zero, self._zero)
        self.n = 0

    def step(self) -> None:
        self.engine.step(*self.macro)
        _, u, v, w = self.macro
        self.transport.step(u, v, w, self._zero, Q=self.Q)
        self.n += 1

    def nusselt(self) -> dict:
        return nusselt(self.transport.populations, self.transport.T, self.flag, self.kappa_f)
