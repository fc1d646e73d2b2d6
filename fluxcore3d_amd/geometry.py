"""Synthetic geometries of the benchmark workloads (seeded, reproducible).

* ``sphere_channel``: 64x32x32 channel, sphere r=6 at (20,16,16), analytic SDF
  |x - c| - r sampled at the nodes (origin 0, pitch 1), walls on +-Y and +-Z.
* ``rsa_spheres``: periodic random sequential addition of non-overlapping
  spheres with radii uniform in [rmin, rmax] until the target solid fraction
  (attempts capped; the achieved porosity is returned), union SDF with the
  minimum-image distance.
* ``gyroid``: periodic gyroid sheet, solid where
  sin x cos y + sin y cos z + sin z cos x < level with x = 2 pi i / period.
* ``pin_fin_plate``: base plate with a square array of cylindrical pins in a
  walled channel (flow along the last axis).

Each returns float32 ``phi`` (negative in solid) and the boolean ``solid``.
"""
from __future__ import annotations

import math

import numpy as np


def sphere_channel(nx=64, ny=32, nz=32, center=(20.0, 16.0, 16.0), radius=6.0, walls=("+Y", "-Y", "+Z", "-Z")):
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    phi = (np.sqrt((i - center[0]) ** 2 + (j - center[1]) ** 2 + (k - center[2]) ** 2) - radius).astype(np.float32)
    solid = phi < 0
    solid_flow = solid.copy()
    sl = {"+X": (slice(-1, None), slice(None), slice(None)), "-X": (slice(0, 1), slice(None), slice(None)),
          "+Y": (slice(None), slice(-1, None), slice(None)), "-Y": (slice(None), slice(0, 1), slice(None)),
          "+Z": (slice(None), slice(None), slice(-1, None)), "-Z": (slice(None), slice(None), slice(0, 1))}
    for face in walls:
        solid_flow[sl[face]] = True
    return phi, solid, solid_flow


def rsa_spheres(n=256, porosity=0.6, rmin=4.0, rmax=10.0, seed=0, tries_per_sphere=64):
    """Random sequential addition of non-overlapping spheres in a periodic box.

    Radii are drawn uniform in [rmin, rmax] until their volume reaches the
    target solid fraction (plus 25 % slack for rejections), then inserted
    largest first at random positions (``tries_per_sphere`` attempts each;
    a sphere that does not fit is dropped). Plain RSA in random order jams
    near a solid fraction of 0.34 for this radius range; inserting by size
    reaches 0.40. Returns phi, solid, info (achieved porosity, counts, seed).
    """
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    L = float(n)
    target = (1.0 - porosity) * L ** 3
    radii_all = []
    v = 0.0
    while v < 1.25 * target:
        r = rng.uniform(rmin, rmax)
        radii_all.append(r)
        v += 4.0 / 3.0 * math.pi * r ** 3
    radii_all = np.sort(np.asarray(radii_all))[::-1]
    centers = np.zeros((0, 3))
    radii = np.zeros(0)
    tree = None
    tree_n = 0
    vol = 0.0
    attempts = 0
    for r in radii_all:
        if vol >= target:
            break
        cand = rng.uniform(0.0, L, size=(tries_per_sphere, 3))
        attempts += tries_per_sphere
        ok = np.ones(tries_per_sphere, bool)
        if tree is not None:
            hits = tree.query_ball_point(cand, r + rmax)
            for t, h in enumerate(hits):
                if h:
                    d = centers[h] - cand[t]
                    d -= L * np.round(d / L)
                    if np.any(np.einsum("ij,ij->i", d, d) < (r + radii[h]) ** 2):
                        ok[t] = False
        recent = slice(tree_n, len(radii))
        if len(radii) > tree_n:
            d = centers[recent][None, :, :] - cand[:, None, :]
            d -= L * np.round(d / L)
            ok &= ~np.any(np.einsum("tij,tij->ti", d, d) < (r + radii[recent][None, :]) ** 2, axis=1)
        good = np.nonzero(ok)[0]
        if good.size == 0:
            continue
        centers = np.vstack([centers, cand[good[0]]])
        radii = np.append(radii, r)
        vol += 4.0 / 3.0 * math.pi * r ** 3
        if len(radii) - tree_n >= 256:
            tree = cKDTree(np.mod(centers, L), boxsize=L)
            tree_n = len(radii)
    phi = np.full((n, n, n), 1e3, dtype=np.float32)
    ar = np.arange(n)
    for c, r in zip(centers, radii):
        m = int(math.ceil(r + 2))
        idx = [np.arange(int(math.floor(c[a])) - m, int(math.floor(c[a])) + m + 2) for a in range(3)]
        dist = np.sqrt((idx[0] - c[0])[:, None, None] ** 2 + (idx[1] - c[1])[None, :, None] ** 2
                       + (idx[2] - c[2])[None, None, :] ** 2) - r
        sel = np.ix_(ar[idx[0] % n], ar[idx[1] % n], ar[idx[2] % n])
        phi[sel] = np.minimum(phi[sel], dist.astype(np.float32))
    solid = phi < 0
    info = {"porosity": float(1.0 - solid.mean()), "n_spheres": int(len(radii)), "attempts": int(attempts),
            "seed": int(seed), "target_porosity": float(porosity)}
    return phi, solid, info


def smooth_blobs(n, seed=0, sigma=None, device="cuda"):
    """Seeded smooth-noise blobs: white noise filtered by a Gaussian of width
    ``sigma`` (default n/16 cells) in Fourier space, thresholded at the median
    (50 % solid). Returns a boolean torch tensor on ``device``."""
    import torch

    g = torch.Generator(device="cpu").manual_seed(int(seed))
    noise = torch.randn((n, n, n), generator=g, dtype=torch.float32).to(device)
    s = float(sigma if sigma is not None else n / 16.0)
    k = torch.fft.fftfreq(n, device=device, dtype=torch.float32)
    kk = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    filt = torch.exp(-2.0 * (torch.pi ** 2) * (s ** 2) * kk)
    field = torch.fft.ifftn(torch.fft.fftn(noise) * filt).real
    thr = torch.quantile(field.flatten()[:: max(1, field.numel() // 2_000_000)].float(), 0.5)
    return field > thr


def conduction_problem(n, seed=0, ratio=None, dt=1.0e4, device="cuda"):
    """Config-5 conduction system A x = b on an n^3 blob field: k = 1 in fluid,
    k = ratio in solid (ratio log-uniform in [1, 1e4] from the seed unless given),
    uniform rho_cp, no flow, Dirichlet x = 0 on the inlet plane (i = 0), Neumann
    elsewhere; b = dt * s with a unit source in the solid, implicit step dt."""
    import numpy as np
    import torch

    rng = np.random.default_rng(seed)
    if ratio is None:
        ratio = float(10.0 ** rng.uniform(0.0, 4.0))
    solid = smooth_blobs(n, seed=seed, device=device)
    k = torch.where(solid, torch.tensor(ratio, device=device), torch.tensor(1.0, device=device)).float()
    mask = torch.ones((n, n, n), dtype=torch.float32, device=device)
    mask[0] = 0.0
    src = solid.float() / max(int(solid.sum().item()), 1)
    b = (dt * src) * mask
    z = torch.zeros_like(k)
    return dict(k=k, rcp=torch.ones_like(k), u=z, v=z, w=z, mask=mask, b=b, dt=dt, ratio=ratio, ax=0, sign=1,
                s=1.0, solid=solid)


def cht_problem(scale=1.0, dt=1.0e3, u_max=0.04, device="cuda"):
    """Conduction-advection system A x = b of the coupled heat-sink configuration
    (pin_fin_plate at 192x192x384 times ``scale``): k_s/k_f = 1000, rho*c_p of the
    solid 2000 times the fluid's, a channel flow along z in the fluid (parabolic in
    x and y, peak ``u_max``, zero in solids and walls; it stands in for the flow
    solution), Dirichlet x = 0 on the inlet plane (k = 0), Neumann elsewhere, and a
    unit heat input spread over the bottom layer of the base plate, times dt. The
    upwind advection makes A non-symmetric."""
    import torch

    nx, ny, nz = int(192 * scale), int(192 * scale), int(384 * scale)
    phi, solid, solid_flow = pin_fin_plate(nx, ny, nz, plate=max(2, int(16 * scale)), radius=4.0 * scale,
                                           height=int(120 * scale))
    S = torch.as_tensor(solid, device=device)
    open_ = ~torch.as_tensor(solid_flow, device=device)
    k = torch.where(S, torch.tensor(1000.0, device=device), torch.tensor(1.0, device=device)).float()
    rcp = torch.where(S, torch.tensor(2000.0, device=device), torch.tensor(1.0, device=device)).float()
    x = torch.linspace(0.0, 1.0, nx, device=device)[:, None, None]
    y = torch.linspace(0.0, 1.0, ny, device=device)[None, :, None]
    w = (u_max * (4 * x * (1 - x)) * (4 * y * (1 - y)) * open_.float()).expand(nx, ny, nz).contiguous()
    zero = torch.zeros_like(k)
    mask = torch.ones_like(k)
    mask[:, :, 0] = 0.0
    src = torch.zeros_like(k)
    src[:, 0, :] = S[:, 0, :].float()
    b = dt * src / src.sum() * mask
    return dict(k=k, rcp=rcp, u=zero, v=zero, w=w, mask=mask, b=b, dt=dt, ax=2, sign=1, s=1.0, solid=S,
                shape=(nx, ny, nz))


def gyroid(n=384, period=48.0, level=0.3, shape=None):
    """Gyroid solid g(x) < level, g = sin x cos y + sin y cos z + sin z cos x,
    x = 2 pi i / period. phi = (g - level) / |grad g| clipped to +-4 (a first-order
    distance in voxels, negative in solid), so Bouzidi q is the trilinear zero
    crossing. With an integer period every period holds the same values bit for
    bit, so a periodic box of whole periods is an exact sub-domain of a larger one."""
    nx, ny, nz = shape if shape is not None else (n, n, n)
    w = 2.0 * math.pi / period

    def coord(m):  # node index mod an integer period: every period gets the same bits
        a = np.arange(m)
        return (a % int(period) if float(period).is_integer() else a) * w

    i, j, k = np.meshgrid(coord(nx), coord(ny), coord(nz), indexing="ij")
    g = np.sin(i) * np.cos(j) + np.sin(j) * np.cos(k) + np.sin(k) * np.cos(i)
    gx = np.cos(i) * np.cos(j) - np.sin(k) * np.sin(i)
    gy = -np.sin(i) * np.sin(j) + np.cos(j) * np.cos(k)
    gz = -np.sin(j) * np.sin(k) + np.cos(k) * np.cos(i)
    grad = w * np.sqrt(gx ** 2 + gy ** 2 + gz ** 2)
    phi = np.clip((g - level) / np.maximum(grad, 1e-3 * w), -4.0, 4.0).astype(np.float32)
    return phi, phi < 0


def pin_fin_plate(nx=192, ny=192, nz=384, plate=16, pins=8, radius=4.0, height=120, margin=None):
    """Heat-sink channel: flow along z, a base plate of ``plate`` voxels at low y,
    a pins x pins array of cylinders (axis along y, radius ``radius``, height
    ``height`` above the plate) centred in x and in the middle third of z.
    Returns phi (SDF of plate union pins, negative in solid), solid and
    solid_flow (solid plus channel walls on +-X and +Y)."""
    i, j, k = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64),
                          np.arange(nz, dtype=np.float64), indexing="ij")
    d_plate = j - (plate - 0.5)
    margin = nx / (2 * pins) if margin is None else margin
    xs = np.linspace(margin, nx - 1 - margin, pins)
    zs = np.linspace(nz / 3.0, 2.0 * nz / 3.0, pins)
    d_pin = np.full(i.shape, np.inf)
    for cx in xs:
        for cz in zs:
            dr = np.sqrt((i - cx) ** 2 + (k - cz) ** 2) - radius
            dh = j - (plate - 0.5 + height)
            d_pin = np.minimum(d_pin, np.maximum(dr, dh))
    phi = np.minimum(d_plate, d_pin).astype(np.float32)
    solid = phi < 0
    solid_flow = solid.copy()
    solid_flow[0, :, :] = solid_flow[-1, :, :] = True
    solid_flow[:, -1, :] = True
    return phi, solid, solid_flow
