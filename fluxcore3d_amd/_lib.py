"""ctypes binding of libfluxcore_hip.so (the C ABI declared in include/fluxcore_hip.h).

The library is loaded from the package directory. There is no fallback: if the
library is missing, ``lib()`` raises (the extension modules then report their
``*_AVAILABLE`` flags as False), and if a GPU call fails, a RuntimeError carries
the message of ``fc_last_error()``.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfluxcore_hip.so")

_lock = threading.Lock()
_lib = None

P = C.c_void_p
I32 = C.c_int
I64 = C.c_int64
F32 = C.c_float
F64 = C.c_double
SZ = C.c_size_t


class AATables(C.Structure):
    _fields_ = [
        ("nflag", P), ("wave_base", P), ("wave_meta", P), ("lk_code", P), ("s_code", P), ("f_code", P),
        ("lane_nodes", P), ("n_lanes", I64), ("sp_mask", P), ("sp_base", P),
        ("w1", P), ("w2", P), ("A", P), ("B", P), ("ff_idx", P), ("n_v", I64),
        ("row_meta", P), ("wpr", I32), ("n_store", I64),
    ]


class AACollision(C.Structure):
    _fields_ = [
        ("kind", I32), ("om1", F32), ("om2", F32), ("S", F32 * 19), ("cs_dx2", F32),
        ("force", F32 * 3), ("use_force", I32), ("sponge", P), ("sponge_axis", I32),
    ]


# name -> (restype, argtypes); every int-returning entry point is checked.
_SIGS = {
    "fc_last_error": (C.c_char_p, []),
    "fc_version": (I32, []),
    # LBM, reference semantics (lbm_ref.hip)
    "fc_lbm_macro_collide_bgk": (I32, [P, P, P, P, P, P, P, F32, I64, P]),
    "fc_lbm_macro_collide_trt": (I32, [P, P, P, P, P, P, P, F32, F32, I64, P]),
    "fc_lbm_macro_collide_mrt": (I32, [P, P, P, P, P, P, P, P, I64, P]),
    "fc_lbm_macro_collide_mrt_smag": (I32, [P, P, P, P, P, P, P, P, F32, F32, I64, P]),
    "fc_lbm_stream_pull": (I32, [P, P, I32, I32, I32, I32, I32, I32, P]),
    "fc_lbm_sponge": (I32, [P, P, P, P, P, P, I64, P]),
    "fc_lbm_ibb": (I32, [P, P, P, P, P, P, P, P, P, P, I64, I64, P]),
    "fc_lbm_reset_solid": (I32, [P, P, I64, P]),
    "fc_lbm_zou_he_bc": (I32, [P, P, I32, I32, I32, I32, I32, I32, I32, F32, F32, I32, P]),
    "fc_lbm_convective_bc_workspace_bytes": (SZ, [I32, I32]),
    "fc_lbm_convective_bc": (I32, [P, P, I32, I32, I32, I32, I32, I32, I32, F32, I32, I32, P, SZ, P]),
    "fc_lbm_convective_bc_noalloc": (I32, [P, P, I32, I32, I32, I32, I32, I32, I32, F32, I32, I32, P]),
    # LBM fused AA step (lbm_aa.hip)
    "fc_aa_slot_count": (I64, [I32, I32, I32, I32, I32, I32]),
    "fc_aa_step": (I32, [P, I32, I32, I32, I32, I32, I32, I32, C.POINTER(AATables), C.POINTER(AACollision),
                         I32, I32, P, P, P, P, I32, P]),
    "fc_aa_materialize": (I32, [P, I32, I32, I32, I32, I32, I32, I32, C.POINTER(AATables), I32, I32, P, P]),
    "fc_aa_load": (I32, [P, I32, I32, I32, I32, I32, I32, I32, C.POINTER(AATables), P, P]),
    "fc_aa_zou_he": (I32, [P, I32, I32, I32, I32, I32, I32, I32, C.POINTER(AATables), I32, I32, I32, I32, I32,
                           I32, I32, F32, F32, I32, P]),
    "fc_aa_convective_workspace_bytes": (SZ, [I32, I32]),
    "fc_aa_convective": (I32, [P, I32, I32, I32, I32, I32, I32, I32, C.POINTER(AATables), I32, I32, I32, I32, I32,
                               I32, I32, F32, P, SZ, P]),
}

_UNCHECKED = {"fc_last_error", "fc_version", "fc_lbm_convective_bc_workspace_bytes", "fc_aa_slot_count",
              "fc_aa_convective_workspace_bytes"}


def register(sigs: dict, unchecked=()) -> None:
    """Add signatures of further entry points (used by the MG/thermal/conv wrappers)."""
    _SIGS.update(sigs)
    _UNCHECKED.update(unchecked)
    if _lib is not None:
        _bind(_lib, sigs)


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """Load libfluxcore_hip.so once; raise if it is absent."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f"libfluxcore_hip.so not found at {LIB_PATH}; build it with "
                        "`python -m fluxcore3d_amd.build` (the GPU path has no fallback)")
                handle = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
                _bind(handle, _SIGS)
                _lib = handle
    return _lib


def load_error() -> str | None:
    """Load the library; return None on success, else the reason it failed."""
    try:
        lib()
    except (RuntimeError, OSError) as e:
        return str(e)
    return None


def exported_symbols() -> list:
    return sorted(_SIGS)


def call(name: str, *args):
    """Call an entry point; non-zero status -> RuntimeError(fc_last_error())."""
    fn = getattr(lib(), name)
    rc = fn(*args)
    if name in _UNCHECKED:
        return rc
    if rc != 0:
        msg = lib().fc_last_error().decode(errors="replace")
        raise RuntimeError(msg or f"{name} failed with status {rc}")
    return rc


def ptr(t) -> int | None:
    """Device pointer of a tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_handle(device=None) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream
