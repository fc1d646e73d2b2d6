"""Shared workload builders for the tests (synthetic geometry, link sets from the oracle)."""
import functools
import types

import numpy as np

from fluxcore3d_amd import geometry
from fluxcore3d_amd.solver.flow_state import FlowState
from fluxcore3d_amd.solver.d3q19 import mrt_rates
from oracle import lbm as olbm


class Cfg(types.SimpleNamespace):
    pass


# the minimal drop-in flow state lives in the package (the benchmark uses it too)
MiniSim = FlowState


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


def gyroid_heat_params():
    """Config 4's conjugate heat leg: MRT flow (tau 0.55, other rates 1.2) driven by a
    body force 1e-5 along x through the period-48 gyroid from rest; fluid kappa =
    nu / 0.71, solid conductivity 10x the fluid's (equal rho c_p), source 1e-4 per
    solid node per step balanced by a uniform sink in the fluid; 5000 steps."""
    tau = 0.55
    om = 1.0 / tau
    return dict(omega=om, S=mrt_rates(om, s_e=1.2, s_bulk=1.2, s_ghost=1.2), force=(1e-5, 0.0, 0.0),
                kappa_f=(tau - 0.5) / 3.0 / 0.71, k_ratio=10.0, q_solid=1e-4, steps=5000)
