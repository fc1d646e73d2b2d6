"""D3Q19 tables in the reference's velocity ordering (LBM_CHT.py:535-558).

W holds the reference's FP32 weight literals (lbm_kernels.cu:65-71), which
solid-node resets reproduce bit for bit.
"""
import numpy as np

C = np.array([
    [0, 0, 0],
    [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [1, -1, 0], [-1, 1, 0], [-1, -1, 0],
    [1, 0, 1], [1, 0, -1], [-1, 0, 1], [-1, 0, -1],
    [0, 1, 1], [0, 1, -1], [0, -1, 1], [0, -1, -1],
], dtype=np.int64)

OPP = np.array([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15], dtype=np.int64)

W = np.array([0.33333333] + [0.05555556] * 6 + [0.02777778] * 12, dtype=np.float32)

# MRT relaxation-rate layout of flow_solver_upgrade.py:216-274
CONSERVED = (0, 3, 5, 7)
VISCOUS = (9, 11, 13, 14, 15)
ENERGY = 1
BULK = (2, 4, 6, 8, 10, 12)
GHOST = (16, 17, 18)


def mrt_rates(omega_nu: float, s_e: float = 1.19, s_bulk: float = 1.4, s_ghost: float = 1.2) -> np.ndarray:
    """19 MRT rates by moment row: conserved 0, viscous omega_nu, energy s_e,
    energy-square / heat-flux / stress-energy s_bulk, ghosts s_ghost."""
    for name, val in (("s_bulk", s_bulk), ("s_ghost", s_ghost), ("s_e", s_e), ("omega_nu", omega_nu)):
        if not (0.0 < val < 2.0):
            raise ValueError(f"MRT relaxation rate {name}={val:.6g} is outside (0,2).")
    S = np.zeros(19, np.float32)
    S[list(VISCOUS)] = omega_nu
    S[ENERGY] = s_e
    S[list(BULK)] = s_bulk
    S[list(GHOST)] = s_ghost
    return S
