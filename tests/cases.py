"""Shared workload builders for the tests (synthetic geometry, link sets from the oracle)."""
import functools
import types

import numpy as np

from fluxcore3d_amd import geometry
from fluxcore3d_amd.solver.d3q19 import W, mrt_rates
from oracle import lbm as olbm


class Cfg(types.SimpleNamespace):
    pass


class MiniSim:
    """The fields of LBMCHT3D_Torch (LBM_CHT.py:562-720) that the drop-in patch reads."""

    def __init__(self, nx, ny, nz, cfg, solid, solid_flow, phi, omega, device, S=None, links=None, omega_minus=None):
        import torch
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
        s = self.solid_flow
        if not s.any():
            return
        f_copy = self.f.clone()
        opp = [0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15]
        for i in range(19):
            self.f[i, s] = f_copy[opp[i], s]


def config1():
    """64x32x32 channel, sphere r=6 at (20,16,16), walls +-Y +-Z, inlet u=0.05 on x=0,
    convective outlet, BGK tau=0.6."""
    phi, solid, solid_flow = geometry.sphere_channel()
    cfg = Cfg(flow_dir="+X", domain_walls=["+Y", "-Y", "+Z", "-Z"], domain_outlets=[], transverse_walls=False,
              outlet_bc_mode="convective", rho_out_lat=1.0, u_lat_max=0.12, collision="bgk", sponge_length=0,
              u_in_lat=0.05)
    links = olbm.ibb_from_sdf(solid, phi)
    return dict(nx=64, ny=32, nz=32, phi=phi, solid=solid, solid_flow=solid_flow, cfg=cfg, links=links,
                omega=1.0 / 0.6)


def pin_fin_channel(nx=48, ny=48, nz=96):
    """Reduced heat-sink channel (the 192x192x384 plate + 8x8 pins workload at
    1/4 scale): plate 4 voxels, 4x4 pins r=2 h=30, flow +Z, inlet u=0.04,
    convective outlet, walls +-X +-Y, TRT with magic parameter 1/4."""
    phi, solid, _ = geometry.pin_fin_plate(nx, ny, nz, plate=4, pins=4, radius=2.0, height=30)
    cfg = Cfg(flow_dir="+Z", domain_walls=["+X", "-X", "+Y", "-Y"], domain_outlets=[], transverse_walls=False,
              outlet_bc_mode="convective", rho_out_lat=1.0, u_lat_max=0.12, collision="trt", sponge_length=0,
              u_in_lat=0.04)
    solid_flow = solid.copy()
    solid_flow[0], solid_flow[-1] = True, True
    solid_flow[:, 0], solid_flow[:, -1] = True, True
    links = olbm.ibb_from_sdf(solid, phi)
    tau = 0.6
    return dict(nx=nx, ny=ny, nz=nz, phi=phi, solid=solid, solid_flow=solid_flow, cfg=cfg, links=links,
                omega=1.0 / tau, omega_minus=1.0 / (0.5 + 0.25 / (tau - 0.5)))


@functools.lru_cache(maxsize=1)
def pin_fin_heat_sink():
    """The coupled heat-sink workload at its full size: 192x192x384 channel, 16-voxel base
    plate with an 8x8 grid of pins (radius 4, height 120), flow +Z, inlet u = 0.04,
    convective outlet, walls +-X +-Y, TRT with magic parameter 1/4."""
    phi, solid, _ = geometry.pin_fin_plate()
    nx, ny, nz = solid.shape
    cfg = Cfg(flow_dir="+Z", domain_walls=["+X", "-X", "+Y", "-Y"], domain_outlets=[], transverse_walls=False,
              outlet_bc_mode="convective", rho_out_lat=1.0, u_lat_max=0.12, collision="trt", sponge_length=0,
              u_in_lat=0.04)
    solid_flow = solid.copy()
    solid_flow[0], solid_flow[-1] = True, True
    solid_flow[:, 0], solid_flow[:, -1] = True, True
    links = olbm.ibb_from_sdf(solid, phi)
    tau = 0.6
    return dict(nx=nx, ny=ny, nz=nz, phi=phi, solid=solid, solid_flow=solid_flow, cfg=cfg, links=links,
                omega=1.0 / tau, omega_minus=1.0 / (0.5 + 0.25 / (tau - 0.5)))


def gyroid_box(n=48, period=48.0):
    """One period of the gyroid (the 384^3, period-48 precision workload's unit cell)."""
    phi, solid = geometry.gyroid(n, period=period)
    links = olbm.ibb_from_sdf(solid, phi, periodic=(True, True, True))
    return dict(nx=n, ny=n, nz=n, phi=phi, solid=solid, solid_flow=solid, links=links)


def periodic_pack(n=48, seed=3):
    phi, solid, info = geometry.rsa_spheres(n, porosity=0.6, rmin=4, rmax=10, seed=seed)
    links = olbm.ibb_from_sdf(solid, phi)
    return dict(nx=n, ny=n, nz=n, phi=phi, solid=solid, solid_flow=solid, links=links, info=info)


def mrt_S(omega):
    return mrt_rates(omega)
