"""Minimal flow state for the drop-in step: the fields of LBMCHT3D_Torch
(LBM_CHT.py:562-720) that LBMCUDAUpgrade.patch reads and writes, for callers that
drive the solver without the reference's simulation class (the coupled
heat-transfer loop, the benchmark).

Populations start at rest (f = w_i, rho = 1, u = 0); ``links`` is an IBBData
link set (pre_post/sdf_tracing.py:8-18).
"""
from __future__ import annotations

import numpy as np
import torch

from .d3q19 import W


class FlowState:
    def __init__(self, nx, ny, nz, cfg, solid, solid_flow, phi, omega, device, S=None, links=None, omega_minus=None):
        self.cfg = cfg
        self.nx, self.ny, self.nz = nx, ny, nz
        self.device = torch.device(device)
        self.dtype = torch.float32
        self.solid = torch.as_tensor(solid, device=self.device)
        self.solid_flow = torch.as_tensor(solid_flow, device=self.device)
        self.phi = torch.as_tensor(phi, device=self.device)
        self.omega_plus = torch.tensor(float(omega))
        self.omega = self.omega_plus
        if omega_minus is not None:
            self.omega_minus = torch.tensor(float(omega_minus))
        self.u_in_lat = float(getattr(cfg, "u_in_lat", 0.05))
        self.f = torch.as_tensor(np.broadcast_to(W[:, None, None, None], (19, nx, ny, nz)).copy(), device=self.device)
        self.rho = torch.ones((nx, ny, nz), device=self.device)
        self.u = torch.zeros((nx, ny, nz), device=self.device)
        self.v = torch.zeros_like(self.u)
        self.wv = torch.zeros_like(self.u)
        if S is not None:
            self._mrt_S = torch.as_tensor(S)
        self.ibb_data = links

    def _bounce_back_post_stream(self):
        """Full-way bounce-back at solid nodes (LBM_CHT.py:1330-1336), used when no
        link set is given."""
        s = self.solid_flow
        if not s.any():
            return
        f_copy = self.f.clone()
        opp = [0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15]
        for i in range(19):
            self.f[i, s] = f_copy[opp[i], s]
