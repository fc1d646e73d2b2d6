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
