"""Drop-in replacement of solver/lbm_cuda_ext.py on MI355X (gfx950).

Same module globals and call surface as the reference (lbm_cuda_ext.py:49-94,
lbm_kernels.cu:1347-1386): ``LBM_CUDA`` exposes macro_collide_{bgk,trt,mrt,
mrt_smag}, stream_pull, sponge, ibb, reset_solid, zou_he_bc and convective_bc
with the reference's argument order and in-place semantics; they run the HIP
kernels of libfluxcore_hip.so through its C ABI. ``LBMCUDAUpgrade.patch(sim)``
replaces ``sim.step`` with the fused in-place AA step (one kernel launch per
step plus the inlet/outlet plane kernels). ``cfg.fc_engine = "reference"``
selects the reference's per-kernel sequence instead.

There is no PyTorch arithmetic inside this module. As in the reference, importing it
never raises: without a GPU, or when libfluxcore_hip.so cannot be loaded,
``LBM_CUDA_AVAILABLE`` is False (and ``LBM_CUDA`` None when the library is missing),
``patch`` leaves the simulation to the caller's own step, and calls raise.
"""
from __future__ import annotations

import sys
import time
import types

import numpy as np
import torch

from .. import _lib
from .aa_engine import AAEngine
from .face_roles import flow_axis_sign, periodic_flags, resolve_face_roles

# ----------------------------------------------------------------------------
# tensor checks (messages follow the reference's TORCH_CHECK texts,
# lbm_kernels.cu:620-624)
# ----------------------------------------------------------------------------
_DT_NAMES = {torch.float32: "float32", torch.uint8: "uint8", torch.int32: "int32", torch.float64: "float64"}


def _check(t, name, dtype, numel=None):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be on CUDA")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if t.dtype != dtype:
        raise RuntimeError(f"{name} must be {_DT_NAMES.get(dtype, dtype)}")
    if numel is not None and t.numel() != numel:
        raise RuntimeError(f"{name} must have {numel} elements (got {t.numel()})")
    return t.data_ptr()


def _stream(t):
    return torch.cuda.current_stream(t.device).cuda_stream


class _LBMKernels:
    """Python face of the LBM entry points (arguments as in lbm_kernels.cu:1347-1386)."""

    def _collide_args(self, f, f_post, rho_out, u_out, v_out, w_out, solid):
        N = f_post.numel() // 19
        ptrs = [_check(f, "f", torch.float32, 19 * N), _check(f_post, "f_post", torch.float32, 19 * N)]
        for t, n in ((rho_out, "rho_out"), (u_out, "u_out"), (v_out, "v_out"), (w_out, "w_out")):
            ptrs.append(_check(t, n, torch.float32, N))
        ptrs.append(_check(solid, "solid", torch.uint8, N))
        return ptrs, N

    def macro_collide_bgk(self, f, f_post, rho_out, u_out, v_out, w_out, solid, omega):
        p, N = self._collide_args(f, f_post, rho_out, u_out, v_out, w_out, solid)
        _lib.call("fc_lbm_macro_collide_bgk", *p, float(omega), N, _stream(f))

    def macro_collide_trt(self, f, f_post, rho_out, u_out, v_out, w_out, solid, omega_plus, omega_minus):
        p, N = self._collide_args(f, f_post, rho_out, u_out, v_out, w_out, solid)
        _lib.call("fc_lbm_macro_collide_trt", *p, float(omega_plus), float(omega_minus), N, _stream(f))

    def macro_collide_mrt(self, f, f_post, rho_out, u_out, v_out, w_out, solid, S_rates):
        p, N = self._collide_args(f, f_post, rho_out, u_out, v_out, w_out, solid)
        s = _check(S_rates, "S_rates", torch.float32)
        if S_rates.numel() != 19:
            raise RuntimeError("S_rates must have 19 elements")
        _lib.call("fc_lbm_macro_collide_mrt", *p, s, N, _stream(f))

    def macro_collide_mrt_smag(self, f, f_post, rho, u, v, w, solid, S_rates, Cs_dx2, omega0):
        p, N = self._collide_args(f, f_post, rho, u, v, w, solid)
        s = _check(S_rates, "S_rates", torch.float32)
        if S_rates.numel() != 19:
            raise RuntimeError("S_rates must have 19 elements")
        _lib.call("fc_lbm_macro_collide_mrt_smag", *p, s, float(Cs_dx2), float(omega0), N, _stream(f))

    def stream_pull(self, f_in, f_out, nx, ny, nz, periodic_x, periodic_y, periodic_z):
        N = int(nx) * int(ny) * int(nz)
        a = _check(f_in, "f_in", torch.float32, 19 * N)
        b = _check(f_out, "f_out", torch.float32, 19 * N)
        _lib.call("fc_lbm_stream_pull", a, b, int(nx), int(ny), int(nz), int(periodic_x), int(periodic_y),
                  int(periodic_z), _stream(f_in))

    def sponge(self, f, rho, u, v, w, sigma):
        N = rho.numel()
        ptrs = [_check(f, "f", torch.float32, 19 * N)]
        for t, n in ((rho, "rho"), (u, "u"), (v, "v"), (w, "w"), (sigma, "sigma")):
            ptrs.append(_check(t, n, torch.float32, N))
        _lib.call("fc_lbm_sponge", *ptrs, N, _stream(f))

    def ibb(self, f_post, f_stream, fluid_flat, dir_i, dir_opp, w1, w2, q_ge_half, ff_flat, ff_valid):
        n = fluid_flat.numel()
        if n == 0:
            return  # lbm_kernels.cu:737
        N = f_post.numel() // 19
        ptrs = [_check(f_post, "f_post", torch.float32, 19 * N), _check(f_stream, "f_stream", torch.float32, 19 * N)]
        for t, nm, dt in ((fluid_flat, "fluid_flat", torch.int32), (dir_i, "dir_i", torch.int32),
                          (dir_opp, "dir_opp", torch.int32), (w1, "w1", torch.float32), (w2, "w2", torch.float32),
                          (q_ge_half, "q_ge_half", torch.uint8), (ff_flat, "ff_flat", torch.int32),
                          (ff_valid, "ff_valid", torch.uint8)):
            ptrs.append(_check(t, nm, dt, n))
        _lib.call("fc_lbm_ibb", *ptrs, n, N, _stream(f_post))

    def reset_solid(self, f, solid):
        N = solid.numel()
        _lib.call("fc_lbm_reset_solid", _check(f, "f", torch.float32, 19 * N), _check(solid, "solid", torch.uint8, N),
                  N, _stream(f))

    def zou_he_bc(self, f, solid, nx, ny, nz, axis, plane_idx, interior_idx, bc_type, u_in_lat, rho_out_lat, sign):
        N = int(nx) * int(ny) * int(nz)
        _lib.call("fc_lbm_zou_he_bc", _check(f, "f", torch.float32, 19 * N), _check(solid, "solid", torch.uint8, N),
                  int(nx), int(ny), int(nz), int(axis), int(plane_idx), int(interior_idx), int(bc_type),
                  float(u_in_lat), float(rho_out_lat), int(sign), _stream(f))

    def convective_bc(self, f, solid, nx, ny, nz, axis, outlet_idx, interior_idx, unknown_cax, rho_out_lat, n1, n2):
        N = int(nx) * int(ny) * int(nz)
        need = int(_lib.call("fc_lbm_convective_bc_workspace_bytes", int(n1), int(n2)))
        ws = torch.empty(max(need, 16), dtype=torch.uint8, device=f.device)
        _lib.call("fc_lbm_convective_bc", _check(f, "f", torch.float32, 19 * N),
                  _check(solid, "solid", torch.uint8, N), int(nx), int(ny), int(nz), int(axis), int(outlet_idx),
                  int(interior_idx), int(unknown_cax), float(rho_out_lat), int(n1), int(n2), ws.data_ptr(),
                  ws.numel(), _stream(f))


def _load_extension() -> bool:
    """Load libfluxcore_hip.so; on failure LBM_CUDA is None and LBM_CUDA_AVAILABLE is
    False, as the reference leaves them when its extension fails (lbm_cuda_ext.py:87-91)."""
    global LBM_CUDA, LBM_CUDA_AVAILABLE
    err = _lib.load_error()
    if err is not None:
        print(f"[LBM_CUDA] libfluxcore_hip.so not loaded: {err}", file=sys.stderr)
        LBM_CUDA, LBM_CUDA_AVAILABLE = None, False
        return False
    LBM_CUDA = _LBMKernels()
    LBM_CUDA_AVAILABLE = bool(torch.cuda.is_available())
    return LBM_CUDA_AVAILABLE


LBM_CUDA = None
LBM_CUDA_AVAILABLE = False
_load_extension()


# ----------------------------------------------------------------------------
# IBB tensor preparation (lbm_cuda_ext.py:100-123)
# ----------------------------------------------------------------------------
def _prepare_ibb_tensors(ibb_data, device, dtype=None):
    if hasattr(ibb_data, "_cuda_tensors"):
        return ibb_data._cuda_tensors

    def _to(t, dt):
        return t.to(device=device, dtype=dt).contiguous()

    t = types.SimpleNamespace(
        fluid_flat=_to(ibb_data.fluid_flat, torch.int32),
        dir_i=_to(ibb_data.dir_i, torch.int32),
        dir_opp=_to(ibb_data.dir_opp, torch.int32),
        w_i=_to(ibb_data.w_i, torch.float32),
        w_second=_to(ibb_data.w_second, torch.float32),
        q_ge_half=_to(ibb_data.q_ge_half.to(torch.uint8), torch.uint8),
        ff_flat=_to(ibb_data.ff_flat, torch.int32),
        ff_valid=_to(ibb_data.ff_valid.to(torch.uint8), torch.uint8),
    )
    ibb_data._cuda_tensors = t
    return t


def _links_numpy(ibb_data):
    if ibb_data is None or int(ibb_data.n_links) == 0:
        return None
    cpu = lambda x: x.detach().cpu().numpy()  # noqa: E731
    return {
        "fluid_flat": cpu(ibb_data.fluid_flat), "dir_i": cpu(ibb_data.dir_i), "w1": cpu(ibb_data.w_i),
        "w2": cpu(ibb_data.w_second), "q_ge_half": cpu(ibb_data.q_ge_half), "ff_flat": cpu(ibb_data.ff_flat),
        "ff_valid": cpu(ibb_data.ff_valid),
    }


# ----------------------------------------------------------------------------
# flow BC plan: the calls _apply_flow_bcs_cuda makes (lbm_cuda_ext.py:628-687)
# ----------------------------------------------------------------------------
def flow_bc_plan(sim) -> list:
    cfg = sim.cfg
    u_max = float(getattr(cfg, "u_lat_max", 0.5))
    rho_out = float(cfg.rho_out_lat)
    flow_axis, flow_sign = flow_axis_sign(cfg)
    u_signed = float(max(-u_max, min(u_max, float(sim.u_in_lat) * flow_sign)))
    outlet_mode = getattr(cfg, "outlet_bc_mode", "pressure").lower()
    dims = (sim.nx, sim.ny, sim.nz)
    plan = []
    for face, role in resolve_face_roles(cfg).items():
        if role not in ("inlet", "outlet"):
            continue
        ax = {"X": 0, "Y": 1, "Z": 2}[face[1]]
        fsign = +1 if face[0] == "+" else -1
        plane = dims[ax] - 1 if fsign == +1 else 0
        interior = plane - 1 if plane > 0 else 1
        if role == "inlet":
            plan.append(dict(kind="zou_he", axis=ax, plane=plane, interior=interior, bc_type=0, u_in=u_signed,
                             rho_out=rho_out, sign=flow_sign))
        elif outlet_mode == "convective" and ax == flow_axis:
            plan.append(dict(kind="convective", axis=ax, plane=plane, interior=interior,
                             unknown_cax=+1 if plane == 0 else -1, rho_out=rho_out))
        else:
            plan.append(dict(kind="zou_he", axis=ax, plane=plane, interior=interior, bc_type=1, u_in=u_signed,
                             rho_out=rho_out, sign=fsign))
    return plan


def _apply_flow_bcs_cuda(sim_self):
    """Per-plane BC calls on sim.f (reference-sequence engine)."""
    N = sim_self.nx * sim_self.ny * sim_self.nz
    f_flat = sim_self.f.reshape(19, N)
    s_flat = sim_self._solid_u8.reshape(N)
    dims = {0: (sim_self.ny, sim_self.nz), 1: (sim_self.nx, sim_self.nz), 2: (sim_self.nx, sim_self.ny)}
    for bc in flow_bc_plan(sim_self):
        if bc["kind"] == "zou_he":
            LBM_CUDA.zou_he_bc(f_flat, s_flat, sim_self.nx, sim_self.ny, sim_self.nz, bc["axis"], bc["plane"],
                               bc["interior"], bc["bc_type"], bc["u_in"], bc["rho_out"], bc["sign"])
        else:
            n1, n2 = dims[bc["axis"]]
            LBM_CUDA.convective_bc(f_flat, s_flat, sim_self.nx, sim_self.ny, sim_self.nz, bc["axis"], bc["plane"],
                                   bc["interior"], bc["unknown_cax"], bc["rho_out"], n1, n2)


def _sponge_profile(sim):
    """The sponge field of FlowSolverUpgrade._init_sponge varies along one axis only;
    the fused kernel takes that 1-D profile."""
    sig = sim._sponge_sigma
    ax = int(getattr(sim, "_sponge_axis", 0))
    n = sig.shape[ax]
    prof = sig.movedim(ax, 0).reshape(n, -1)
    if not torch.equal(prof, prof[:, :1].expand_as(prof)):
        raise RuntimeError("sponge sigma must be constant over planes normal to the sponge axis")
    return prof[:, 0].float().cpu().numpy(), ax


# ----------------------------------------------------------------------------
# LBMCUDAUpgrade
# ----------------------------------------------------------------------------
class _FusedF:
    """``sim.f`` of a patched simulation: materialised from the AA state on read,
    loaded into it on assignment."""

    def __get__(self, obj, owner=None):
        if obj is None:
            return self
        eng = obj.__dict__.get("_fc_engine")
        if eng is None:
            return obj.__dict__["f"]
        if obj.__dict__.get("_fc_f_tick") != eng.t:
            buf = obj.__dict__.get("_fc_f_buf")
            if buf is None:
                buf = torch.empty((19, obj.nx, obj.ny, obj.nz), dtype=torch.float32, device=eng.device)
                obj.__dict__["_fc_f_buf"] = buf
            eng.materialize(buf)
            obj.__dict__["_fc_f_tick"] = eng.t
        return obj.__dict__["_fc_f_buf"]

    def __set__(self, obj, value):
        eng = obj.__dict__.get("_fc_engine")
        if eng is None:
            obj.__dict__["f"] = value
            return
        eng.load(value.reshape(19, -1).float())
        obj.__dict__["_fc_f_tick"] = None


class LBMCUDAUpgrade:
    """Patch a simulation object in place (call after FlowSolverUpgrade.patch)."""

    @staticmethod
    def patch(sim):
        if not LBM_CUDA_AVAILABLE:
            print("[LBM_CUDA] Kernels not available -- patch skipped.")
            return
        cfg = sim.cfg
        sim._solid_u8 = sim.solid_flow.to(dtype=torch.uint8, device=sim.device).contiguous()
        mode = getattr(cfg, "collision", "bgk").lower()
        sim._S_cuda = (sim._mrt_S.to(device=sim.device, dtype=torch.float32).contiguous()
                       if mode in ("mrt", "mrt_smag") else None)
        if mode == "mrt_smag":
            sim._cuda_Cs_dx2 = float(float(getattr(cfg, "smag_Cs", 0.1)) ** 2)
            sim._cuda_omega0 = float(sim.omega_plus)
        else:
            sim._cuda_Cs_dx2 = None
            sim._cuda_omega0 = None
        sim._cuda_periodic = periodic_flags(cfg)
        has_sponge = int(getattr(cfg, "sponge_length", 0)) > 0 and hasattr(sim, "_sponge_sigma")
        sim._sponge_sigma_flat = sim._sponge_sigma.reshape(-1).contiguous() if has_sponge else None
        sim._cuda_debug_nan = bool(getattr(cfg, "cuda_debug_nan", False))
        engine = getattr(cfg, "fc_engine", "aa").lower()
        if engine == "reference":
            LBMCUDAUpgrade._patch_reference(sim, mode, has_sponge)
        elif engine == "aa":
            LBMCUDAUpgrade._patch_aa(sim, mode, has_sponge)
        else:
            raise ValueError("cfg.fc_engine must be 'aa' or 'reference'")
        print(f"[LBM_CUDA] LBMCUDAUpgrade patched: engine={engine} collision={mode} "
              f"periodic={sim._cuda_periodic} links={getattr(getattr(sim, 'ibb_data', None), 'n_links', 0)}")

    # ---- fused AA engine ------------------------------------------------------
    @staticmethod
    def _patch_aa(sim, mode, has_sponge):
        cfg = sim.cfg
        omega_minus = float(getattr(sim, "omega_minus", sim.omega_plus)) if mode == "trt" else 0.0
        om1 = {"bgk": float(sim.omega), "trt": float(sim.omega_plus)}.get(mode, float(sim.omega_plus))
        sponge, sponge_axis = (None, 0)
        if has_sponge:
            sponge, sponge_axis = _sponge_profile(sim)
        force = getattr(cfg, "body_force_lat", None)
        eng = AAEngine(sim.nx, sim.ny, sim.nz, sim.solid_flow.detach().cpu().numpy(), periodic=sim._cuda_periodic,
                       collision=mode, omega=om1, omega_minus=omega_minus,
                       S=None if sim._S_cuda is None else sim._S_cuda.cpu().numpy(),
                       cs_dx2=sim._cuda_Cs_dx2 or 0.0, force=force, sponge=sponge, sponge_axis=sponge_axis,
                       links=_links_numpy(getattr(sim, "ibb_data", None)),
                       ibb_mode=getattr(cfg, "ibb_mode", "reference"), bcs=flow_bc_plan(sim),
                       storage=getattr(cfg, "population_storage", "fp32"), device=sim.device,
                       layout=getattr(cfg, "population_layout", "auto"))
        eng.load(sim.f.reshape(19, -1).float())
        for name in ("rho", "u", "v", "wv"):
            t = getattr(sim, name)
            if t.dtype != torch.float32 or not t.is_contiguous():
                setattr(sim, name, t.float().contiguous())
        cls = sim.__class__
        if not getattr(cls, "_fc_fused", False):
            sim.__class__ = type(cls.__name__, (cls,), {"f": _FusedF(), "_fc_fused": True})
        sim.__dict__["_fc_engine"] = eng
        sim._original_step_thermal = None
        sim._macroscopic_python = getattr(sim, "_macroscopic", None)
        sim._macroscopic = types.MethodType(lambda self: None, sim)
        sim.step = types.MethodType(LBMCUDAUpgrade._step_aa, sim)

    @staticmethod
    def _step_aa(sim_self, do_thermal: bool = True):
        eng = sim_self.__dict__["_fc_engine"]
        eng.step(sim_self.rho.view(-1), sim_self.u.view(-1), sim_self.v.view(-1), sim_self.wv.view(-1))
        if sim_self._cuda_debug_nan:
            f = sim_self.f
            if not torch.isfinite(f).all():
                bad = torch.nonzero(~torch.isfinite(f), as_tuple=False)[0]
                print("[LBM_CUDA NaN] after BCs: (q,x,y,z)=", tuple(int(v) for v in bad.tolist()))
                raise FloatingPointError("Non-finite f after BCs")
        return (0, 0)

    # ---- reference kernel sequence -------------------------------------------
    @staticmethod
    def _patch_reference(sim, mode, has_sponge):
        if not hasattr(sim, "_f_post") or sim._f_post.shape != sim.f.shape:
            sim._f_post = torch.empty_like(sim.f)
        sim._f_stream = None
        sim._macroscopic_python = getattr(sim, "_macroscopic", None)
        sim._macroscopic = types.MethodType(lambda self: None, sim)
        sim._collide = types.MethodType(LBMCUDAUpgrade._make_collide(mode), sim)
        sim._stream = types.MethodType(LBMCUDAUpgrade._stream_cuda, sim)
        if getattr(sim, "ibb_data", None) is not None and sim.ibb_data.n_links > 0:
            _prepare_ibb_tensors(sim.ibb_data, sim.device)
            sim._interp_bounce_back_from_post = types.MethodType(LBMCUDAUpgrade._ibb_cuda, sim)
        if has_sponge:
            sim._apply_sponge_zone = types.MethodType(LBMCUDAUpgrade._sponge_cuda, sim)
        sim._apply_flow_bcs = types.MethodType(_apply_flow_bcs_cuda, sim)
        sim._original_step_thermal = getattr(sim, "step", None)
        sim.step = types.MethodType(LBMCUDAUpgrade._step_cuda, sim)

    @staticmethod
    def _make_collide(mode: str):
        def _run(sim_self):
            N = sim_self.nx * sim_self.ny * sim_self.nz
            f_flat = sim_self.f.reshape(19, N).contiguous()
            fp = sim_self._f_post.reshape(19, N)
            args = (f_flat, fp, sim_self.rho.reshape(N), sim_self.u.reshape(N), sim_self.v.reshape(N),
                    sim_self.wv.reshape(N), sim_self._solid_u8.reshape(N))
            if mode == "trt":
                LBM_CUDA.macro_collide_trt(*args, float(sim_self.omega_plus), float(sim_self.omega_minus))
            elif mode == "mrt":
                LBM_CUDA.macro_collide_mrt(*args, sim_self._S_cuda)
            elif mode == "mrt_smag":
                LBM_CUDA.macro_collide_mrt_smag(*args, sim_self._S_cuda, float(sim_self._cuda_Cs_dx2),
                                                float(sim_self._cuda_omega0))
            else:
                LBM_CUDA.macro_collide_bgk(*args, float(sim_self.omega))
            # zero velocity inside solids (lbm_cuda_ext.py:308-310)
            sim_self.u[sim_self.solid_flow] = 0.
            sim_self.v[sim_self.solid_flow] = 0.
            sim_self.wv[sim_self.solid_flow] = 0.
        return _run

    @staticmethod
    def _stream_cuda(sim_self):
        px, py, pz = sim_self._cuda_periodic
        LBM_CUDA.stream_pull(sim_self._f_post.contiguous(), sim_self.f, sim_self.nx, sim_self.ny, sim_self.nz,
                             int(px), int(py), int(pz))

    @staticmethod
    def _ibb_cuda(sim_self, f_post):
        ibb = sim_self.ibb_data
        if ibb.n_links == 0:
            return
        t = ibb._cuda_tensors
        N = sim_self.nx * sim_self.ny * sim_self.nz
        LBM_CUDA.ibb(f_post.reshape(19, N).contiguous(), sim_self.f.reshape(19, N), t.fluid_flat, t.dir_i,
                     t.dir_opp, t.w_i, t.w_second, t.q_ge_half, t.ff_flat, t.ff_valid)
        LBM_CUDA.reset_solid(sim_self.f.reshape(19, N), sim_self._solid_u8.reshape(N))

    @staticmethod
    def _sponge_cuda(sim_self, target_f=None):
        if sim_self._sponge_sigma_flat is None:
            return
        N = sim_self.nx * sim_self.ny * sim_self.nz
        f_tgt = (target_f if target_f is not None else sim_self.f).reshape(19, N)
        LBM_CUDA.sponge(f_tgt, sim_self.rho.reshape(N), sim_self.u.reshape(N), sim_self.v.reshape(N),
                        sim_self.wv.reshape(N), sim_self._sponge_sigma_flat)

    @staticmethod
    def _step_cuda(sim_self, do_thermal: bool = True):
        """collide -> sponge -> stream -> IBB (+reset) or bounce-back -> flow BCs
        (lbm_cuda_ext.py:445-517)."""
        sim_self._collide()
        if sim_self._sponge_sigma_flat is not None:
            sim_self._apply_sponge_zone(sim_self._f_post)
        sim_self._stream()
        if getattr(sim_self, "ibb_data", None) is not None and sim_self.ibb_data.n_links > 0:
            sim_self._interp_bounce_back_from_post(sim_self._f_post)
        else:
            sim_self._bounce_back_post_stream()
        sim_self._apply_flow_bcs()
        if sim_self._cuda_debug_nan and not torch.isfinite(sim_self.f).all():
            bad = torch.nonzero(~torch.isfinite(sim_self.f), as_tuple=False)[0]
            print("[LBM_CUDA NaN] after BCs: (q,x,y,z)=", tuple(int(v) for v in bad.tolist()))
            raise FloatingPointError("Non-finite f after BCs")
        if do_thermal and callable(getattr(sim_self, "_original_step_thermal", None)):
            return sim_self._original_step_thermal()
        return (0, 0)


def add_step_profiler(sim, print_every: int = 200):
    """Synchronising per-phase step timer (lbm_cuda_ext.py:693-748), injected after
    LBMCUDAUpgrade.patch. With the reference kernel sequence the phases are the
    reference's (collide, sponge, stream, ibb, bc); the fused step has two launches
    per step, the fused collide/stream/Bouzidi kernel ("fused") and the inlet/outlet
    plane kernels ("bc"). Prints ms per phase and steps per second every
    ``print_every`` steps."""
    fused = "_fc_engine" in sim.__dict__
    phases = ("fused", "bc") if fused else ("collide", "sponge", "stream", "ibb", "bc")
    sim._prof = {k: 0.0 for k in phases + ("total",)}
    sim._prof["n"] = 0

    def _sync_time():
        torch.cuda.synchronize()
        return time.perf_counter()

    def _timed_step(self, do_thermal=True):
        t = [_sync_time()]
        if fused:
            eng = self.__dict__["_fc_engine"]
            eng.step_fused(self.rho.view(-1), self.u.view(-1), self.v.view(-1), self.wv.view(-1))
            t.append(_sync_time())
            eng.apply_bcs()
            t.append(_sync_time())
        else:
            self._collide()
            t.append(_sync_time())
            if self._sponge_sigma_flat is not None:
                self._apply_sponge_zone(self._f_post)
            t.append(_sync_time())
            self._stream()
            t.append(_sync_time())
            if getattr(self, "ibb_data", None) is not None and self.ibb_data.n_links > 0:
                self._interp_bounce_back_from_post(self._f_post)
            else:
                self._bounce_back_post_stream()
            t.append(_sync_time())
            self._apply_flow_bcs()
            t.append(_sync_time())
        p = self._prof
        for k, a, b in zip(phases, t[:-1], t[1:]):
            p[k] += b - a
        p["total"] += t[-1] - t[0]
        p["n"] += 1
        if p["n"] % print_every == 0:
            n = p["n"]
            print(f"[PROFILE] {n} steps  (ms/step)")
            for k in phases + ("total",):
                print(f"  {k:10s}: {p[k] / n * 1e3:.3f} ms")
            print(f"  it/s      : {n / p['total']:.1f}")
        if not fused and do_thermal and callable(getattr(self, "_original_step_thermal", None)):
            return self._original_step_thermal()
        return (0, 0)

    sim.step = types.MethodType(_timed_step, sim)
    print(f"[PROFILER] Injected. Prints every {print_every} steps.")


__all__ = ["LBM_CUDA", "LBM_CUDA_AVAILABLE", "LBMCUDAUpgrade", "_prepare_ibb_tensors", "_apply_flow_bcs_cuda",
           "add_step_profiler", "flow_bc_plan", "np"]
