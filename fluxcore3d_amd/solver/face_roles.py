"""Face roles of the six domain faces (inlet / outlet / wall / periodic).

Same rule as pre_post/face_roles.py:3-44 of the reference: the inlet face is
the upstream face of ``cfg.flow_dir`` and is never overridden; a user wall
beats the default outlet; the downstream face is the outlet; extra outlets
come from ``cfg.domain_outlets``; ``cfg.transverse_walls`` turns the remaining
faces into walls; everything else is periodic.
"""
from __future__ import annotations

import warnings

ALL_FACES = ("+X", "-X", "+Y", "-Y", "+Z", "-Z")
AXIS = {"X": 0, "Y": 1, "Z": 2}


def flow_axis_sign(cfg) -> tuple[int, int]:
    fd = cfg.flow_dir.upper().strip()
    return AXIS[fd[1]], (+1 if fd[0] == "+" else -1)


def resolve_face_roles(cfg) -> dict:
    axis, sign = flow_axis_sign(cfg)
    ch = "XYZ"[axis]
    inlet = ("-" if sign == +1 else "+") + ch
    outlet = ("+" if sign == +1 else "-") + ch
    walls = {s.upper().strip() for s in getattr(cfg, "domain_walls", [])}
    outlets = {s.upper().strip() for s in getattr(cfg, "domain_outlets", [])}
    legacy = bool(getattr(cfg, "transverse_walls", False))
    roles = {}
    for face in ALL_FACES:
        if face == inlet:
            roles[face] = "inlet"
        elif face in walls:
            roles[face] = "wall"
        elif face == outlet or face in outlets:
            roles[face] = "outlet"
        elif legacy:
            roles[face] = "wall"
        else:
            roles[face] = "periodic"
    if inlet in walls:
        warnings.warn(f"resolve_face_roles: '{inlet}' is the inlet face — cannot be walled. Override ignored.")
    if inlet in outlets:
        warnings.warn(f"resolve_face_roles: '{inlet}' is the inlet face — cannot be set as outlet. Override ignored.")
    return roles


def periodic_flags(cfg) -> tuple[bool, bool, bool]:
    """Streaming periodicity per axis: an axis with any inlet/outlet/wall face is not periodic
    (lbm_cuda_ext.py:197-205)."""
    non_periodic = set()
    for face, role in resolve_face_roles(cfg).items():
        if role in ("inlet", "outlet", "wall"):
            non_periodic.add(AXIS[face[1]])
    return tuple(d not in non_periodic for d in range(3))


def face_ax_idx_fsign(dims, face: str) -> tuple[int, int, int]:
    ax = AXIS[face[1]]
    n = dims[ax]
    return (ax, n - 1, +1) if face[0] == "+" else (ax, 0, -1)
