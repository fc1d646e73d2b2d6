"""Drop-in replacement of solver/mg_cuda_ext.py on MI355X.

``MG_CUDA.mop_fused(T, aE, aW, aN, aS, aU, aD, diagL, Lo, dt, ax, out_idx) -> Lo``
and ``MG_CUDA.jacobi_smooth(x, b, Lo, inv_diagM)`` keep the reference's
arguments and in-place semantics (mg_kernels.cu:133-205); the fused multigrid
entry points used by nlevel_mg_cuda.py are registered here as well.
"""
from __future__ import annotations

import sys

import torch

from .. import _lib

P, I32, I64, F32, F64 = _lib.P, _lib.I32, _lib.I64, _lib.F32, _lib.F64

_lib.register({
    "fc_mg_mop_fused": (I32, [P, P, P, P, P, P, P, P, P, F64, I32, I32, I32, I32, I32, P]),
    "fc_mg_jacobi_smooth": (I32, [P, P, P, P, I64, P]),
    "fc_mg_build_coeffs": (I32, [P, P, P, P, P, I32, I32, I32, I32, I32, F64, F64, P, P, P, P, P, P, P, P]),
    "fc_mg_build_faces": (I32, [P, I32, I32, I32, I32, I32, P, P, P, P, P]),
    "fc_mg_inv_diag": (I32, [P, P, F64, F64, P, I64, P]),
    "fc_mg_restrict_field": (I32, [P, P, I32, I32, I32, I32, P]),
    "fc_mg_sweep": (I32, [P, P, P, P, F64, I32, I32, I32, I32, I32, I32, P]),
    "fc_mg_sweep2_from_zero": (I32, [P, P, P, F64, I32, I32, I32, I32, I32, P]),
    "fc_mg_residual_restrict": (I32, [P, P, P, P, F64, I32, I32, I32, I32, I32, P, P]),
    "fc_mg_prolong_correct": (I32, [P, P, P, I32, I32, I32, I32, I32, I32, F32, P, P]),
    "fc_mg_coarse_solve": (I32, [P, P, P, P, F64, I32, I32, I32, I32, I32, I32, I32, P]),
    "fc_mg_partials_count": (I32, [I64]),
    "fc_mg_apply_A64": (I32, [P, P, P, P, P, F64, I32, I32, I32, I32, I32, P, P, P]),
    "fc_mg_update_xr": (I32, [P, P, P, P, P, I64, P, I32, P, P, P]),
    "fc_mg_new_direction": (I32, [P, P, P, I64, I32, P, I32, P, P, P]),
    "fc_mg_true_residual": (I32, [P, P, P, P, P, P, F64, I32, I32, I32, I32, I32, P, P, P]),
    "fc_mg_norm2_f32": (I32, [P, I64, P, P, I32, P]),
    "fc_mg_residual_f32": (I32, [P, P, P, P, I64, P, P, I32, P]),
    "fc_mg_scalar_slots": (I32, []),
}, unchecked={"fc_mg_partials_count", "fc_mg_scalar_slots"})


def _check(t, name):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be float32")
    return t.data_ptr()


class _MGKernels:
    def mop_fused(self, T, aE, aW, aN, aS, aU, aD, diagL, Lo, dt, ax, out_idx):
        if T.dim() != 3:
            raise RuntimeError("T must be a 3-D (nx, ny, nz) tensor")
        nx, ny, nz = T.shape
        ptrs = [_check(T, "T")]
        for t, n in ((aE, "aE"), (aW, "aW"), (aN, "aN"), (aS, "aS"), (aU, "aU"), (aD, "aD"), (diagL, "diagL"),
                     (Lo, "Lo")):
            ptrs.append(_check(t, n))
            if t.numel() != T.numel():
                raise RuntimeError(f"{n} must have the shape of T")
        _lib.call("fc_mg_mop_fused", *ptrs, float(dt), nx, ny, nz, int(ax), int(out_idx),
                  torch.cuda.current_stream(T.device).cuda_stream)
        return Lo

    def jacobi_smooth(self, x, b, Lo, inv_diagM):
        ptrs = [_check(x, "x"), _check(b, "b"), _check(Lo, "Lo"), _check(inv_diagM, "inv_diagM")]
        for t in (b, Lo, inv_diagM):
            if t.numel() != x.numel():
                raise RuntimeError("x, b, Lo and inv_diagM must have the same size")
        _lib.call("fc_mg_jacobi_smooth", *ptrs, x.numel(), torch.cuda.current_stream(x.device).cuda_stream)


# As in the reference (mg_cuda_ext.py:59-62), a library that cannot be loaded leaves
# MG_CUDA None and MG_CUDA_AVAILABLE False instead of failing the import.
_err = _lib.load_error()
if _err is not None:
    print(f"[MG_CUDA] libfluxcore_hip.so not loaded: {_err}", file=sys.stderr)
    MG_CUDA = None
    MG_CUDA_AVAILABLE = False
else:
    MG_CUDA = _MGKernels()
    MG_CUDA_AVAILABLE = bool(torch.cuda.is_available())
