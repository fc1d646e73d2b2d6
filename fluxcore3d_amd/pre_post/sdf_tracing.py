"""Bouzidi link set and wall distances on the GPU (drop-in for pre_post/sdf_tracing.py).

``build_links_from_mask_gpu``, ``sdf_trilinear``, ``bfl_q_from_sdf``,
``bfl_weights_from_q`` and ``build_ibb_from_sdf_gpu`` keep the reference's
names, arguments and results (sdf_tracing.py:8-259); the per-node
classification, the SDF sampling and the q search run in ``csrc/links.hip``.
The two host-side SDF builders the reference's drivers import from this module
(``build_phi_from_stl_mm``, ``build_phi_from_voxel_mask_mm``,
sdf_tracing.py:262-341) are provided too.
Link order is the reference's: direction q = 1..18 outer, C order inside.
``periodic`` (a build extension, default off like the reference) also links
across periodic faces.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .. import _lib

_lib.register({
    "fc_sdf_trilinear": (_lib.I32, [_lib.P, _lib.I32, _lib.I32, _lib.I32, _lib.F32, _lib.F32, _lib.F32, _lib.F32,
                                    _lib.P, _lib.I64, _lib.P, _lib.P]),
    "fc_bfl_q_rays": (_lib.I32, [_lib.P, _lib.I32, _lib.I32, _lib.I32, _lib.F32, _lib.F32, _lib.F32, _lib.F32,
                                 _lib.P, _lib.P, _lib.P, _lib.I64, _lib.I32, _lib.I32, _lib.P, _lib.P]),
    "fc_bfl_weights": (_lib.I32, [_lib.P, _lib.I64, _lib.P, _lib.P, _lib.P, _lib.P]),
    "fc_links_mask": (_lib.I32, [_lib.P, _lib.I32, _lib.I32, _lib.I32, _lib.I32, _lib.I32, _lib.I32, _lib.P, _lib.P,
                                 _lib.P]),
    "fc_links_q": (_lib.I32, [_lib.P, _lib.I32, _lib.I32, _lib.I32, _lib.F32, _lib.F32, _lib.F32, _lib.F32, _lib.I32,
                              _lib.I32, _lib.I32, _lib.P, _lib.P, _lib.I64, _lib.I32, _lib.I32, _lib.P, _lib.P,
                              _lib.P, _lib.P, _lib.P]),
})

_C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
               [1, -1, 0], [-1, 1, 0], [-1, -1, 0], [1, 0, 1], [1, 0, -1], [-1, 0, 1], [-1, 0, -1], [0, 1, 1],
               [0, 1, -1], [0, -1, 1], [0, -1, -1]])
_OPP = np.array([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15])


@dataclass
class IBBData:
    n_links: int
    fluid_flat: torch.Tensor
    dir_i: torch.Tensor
    dir_opp: torch.Tensor
    q_ge_half: torch.Tensor
    w_i: torch.Tensor
    w_second: torch.Tensor
    ff_flat: torch.Tensor
    ff_valid: torch.Tensor


def build_links_from_mask_gpu(solid_geom: torch.Tensor, c_np=_C, opp_np=_OPP, periodic=(False, False, False)):
    """(fluid_flat, dir_i, dir_opp, ff_flat, ff_valid) of all fluid->solid links."""
    if not solid_geom.is_cuda:
        raise RuntimeError("solid_geom must be on CUDA")
    if not np.array_equal(np.asarray(c_np), _C) or not np.array_equal(np.asarray(opp_np), _OPP):
        raise ValueError("only the D3Q19 ordering of LBM_CHT.py:535-558 is supported")
    nx, ny, nz = solid_geom.shape
    dev = solid_geom.device
    s8 = solid_geom.to(torch.uint8).contiguous()
    mask = torch.empty(nx * ny * nz, dtype=torch.int32, device=dev)
    ffok = torch.empty_like(mask)
    _lib.call("fc_links_mask", s8.data_ptr(), nx, ny, nz, *[int(p) for p in periodic], mask.data_ptr(),
              ffok.data_ptr(), _lib.stream_handle(dev))
    flats, dirs, ffs, oks = [], [], [], []
    for q in range(1, 19):
        sel = torch.nonzero((mask >> q) & 1, as_tuple=False).view(-1)
        if sel.numel() == 0:
            continue
        cx, cy, cz = (int(v) for v in _C[q])
        i, j, k = sel // (ny * nz), (sel // nz) % ny, sel % nz
        iff, jff, kff = i - cx, j - cy, k - cz
        if periodic[0]:
            iff = iff % nx
        if periodic[1]:
            jff = jff % ny
        if periodic[2]:
            kff = kff % nz
        ff = (iff.clamp(0, nx - 1) * ny + jff.clamp(0, ny - 1)) * nz + kff.clamp(0, nz - 1)
        flats.append(sel)
        dirs.append(torch.full_like(sel, q))
        ffs.append(ff)
        oks.append(((ffok[sel] >> q) & 1).bool())
    if not flats:
        z = torch.zeros(0, dtype=torch.int64, device=dev)
        return z, z, z, z, torch.zeros(0, dtype=torch.bool, device=dev)
    fluid_flat = torch.cat(flats).to(torch.int64)
    dir_i = torch.cat(dirs).to(torch.int64)
    dir_opp = torch.as_tensor(_OPP, device=dev)[dir_i]
    return fluid_flat, dir_i, dir_opp, torch.cat(ffs).to(torch.int64), torch.cat(oks)


def bfl_q_and_weights(phi, origin, pitch, fluid_flat, dir_i, n_coarse=6, n_bisect=10,
                      periodic=(False, False, False)):
    """q, q>=1/2, w1, w2 for each link (sdf_tracing.py:127-205); ``periodic`` wraps
    the SDF samples on those axes instead of clamping them (build extension)."""
    phi = phi.to(torch.float32).contiguous()
    nx, ny, nz = phi.shape
    n = int(fluid_flat.numel())
    dev = phi.device
    q = torch.empty(n, dtype=torch.float32, device=dev)
    w1 = torch.empty_like(q)
    w2 = torch.empty_like(q)
    ge = torch.empty(n, dtype=torch.uint8, device=dev)
    org = [float(v) for v in origin]
    ff = fluid_flat.to(torch.int64).contiguous()
    di = dir_i.to(torch.int64).contiguous()
    _lib.call("fc_links_q", phi.data_ptr(), nx, ny, nz, org[0], org[1], org[2], float(pitch),
              *[int(bool(p)) for p in periodic], ff.data_ptr(),
              di.data_ptr(), n, int(n_coarse), int(n_bisect), q.data_ptr(), w1.data_ptr(), w2.data_ptr(),
              ge.data_ptr(), _lib.stream_handle(dev))
    return q, ge.bool(), w1, w2


def _f32_dev(t, dev):
    return torch.as_tensor(t).to(device=dev, dtype=torch.float32).contiguous()


def _origin3(origin):
    o = origin.detach().cpu().reshape(-1) if isinstance(origin, torch.Tensor) else np.asarray(origin).reshape(-1)
    return [float(np.float32(v)) for v in o[:3]]


def sdf_trilinear(phi: torch.Tensor, origin, dx: float, x: torch.Tensor) -> torch.Tensor:
    """Trilinear SDF value at points x (..., 3) in world units (sdf_tracing.py:88-125):
    grid coordinates (x - origin) / dx clamped to [0, n - 1.001], lerps along x, y, z."""
    phi = phi.to(torch.float32).contiguous()
    if not phi.is_cuda:
        raise RuntimeError("phi must be on CUDA")
    nx, ny, nz = phi.shape
    pts = _f32_dev(x, phi.device).reshape(-1, 3)
    out = torch.empty(pts.shape[0], dtype=torch.float32, device=phi.device)
    _lib.call("fc_sdf_trilinear", phi.data_ptr(), nx, ny, nz, *_origin3(origin), float(dx), pts.data_ptr(),
              pts.shape[0], out.data_ptr(), _lib.stream_handle(phi.device))
    return out.reshape(tuple(x.shape[:-1]))


def bfl_q_from_sdf(phi: torch.Tensor, origin, dx: float, x0: torch.Tensor, d: torch.Tensor, L: torch.Tensor,
                   n_coarse: int = 6, n_bisect: int = 10) -> torch.Tensor:
    """Wall distance q in (0, 1] along each ray x0 + t d, t in [0, L] (sdf_tracing.py:127-184):
    first + to - crossing of n_coarse + 1 samples, n_bisect bisections, q = hi / (L + 1e-30);
    0.5 where no crossing is found."""
    phi = phi.to(torch.float32).contiguous()
    if not phi.is_cuda:
        raise RuntimeError("phi must be on CUDA")
    nx, ny, nz = phi.shape
    dev = phi.device
    x0 = _f32_dev(x0, dev).reshape(-1, 3)
    d = _f32_dev(d, dev).reshape(-1, 3)
    L = _f32_dev(L, dev).reshape(-1)
    n = x0.shape[0]
    if d.shape[0] != n or L.shape[0] != n:
        raise RuntimeError("x0, d and L must describe the same number of rays")
    q = torch.empty(n, dtype=torch.float32, device=dev)
    _lib.call("fc_bfl_q_rays", phi.data_ptr(), nx, ny, nz, *_origin3(origin), float(dx), x0.data_ptr(),
              d.data_ptr(), L.data_ptr(), n, int(n_coarse), int(n_bisect), q.data_ptr(), _lib.stream_handle(dev))
    return q


def bfl_weights_from_q(q: torch.Tensor):
    """(q >= 1/2, w1, w2) of the BFL rule (sdf_tracing.py:187-205): 1/(2q) and 1 - 1/(2q)
    for q >= 1/2, 2q and 1 - 2q below."""
    if not q.is_cuda:
        raise RuntimeError("q must be on CUDA")
    qf = q.to(torch.float32).contiguous()
    n = qf.numel()
    ge = torch.empty(qf.shape, dtype=torch.uint8, device=q.device)
    w1 = torch.empty_like(qf)
    w2 = torch.empty_like(qf)
    _lib.call("fc_bfl_weights", qf.data_ptr(), n, ge.data_ptr(), w1.data_ptr(), w2.data_ptr(),
              _lib.stream_handle(q.device))
    return ge.bool(), w1, w2


def build_ibb_from_sdf_gpu(sim, voxel_origin_mm, pitch_mm, offset_ijk, c_np=_C, opp_np=_OPP,
                           periodic=(False, False, False)):
    """Sets ``sim.ibb_data`` from ``sim.solid`` and ``sim.phi`` (sdf_tracing.py:208-259)."""
    origin = (np.asarray(voxel_origin_mm, np.float32)
              - np.float32(pitch_mm) * np.asarray(offset_ijk, np.float32)).astype(np.float32)
    fluid_flat, dir_i, dir_opp, ff_flat, ff_valid = build_links_from_mask_gpu(sim.solid, c_np, opp_np, periodic)
    q, ge, w1, w2 = bfl_q_and_weights(sim.phi, origin, pitch_mm, fluid_flat, dir_i, periodic=periodic)
    sim.ibb_data = IBBData(n_links=int(fluid_flat.numel()), fluid_flat=fluid_flat, dir_i=dir_i, dir_opp=dir_opp,
                           q_ge_half=ge, w_i=w1, w_second=w2, ff_flat=ff_flat, ff_valid=ff_valid)
    sim.ibb_q = q
    return sim.ibb_data


# ----------------------------------------------------------------------------
# Host-side SDF builders (sdf_tracing.py:262-341). They run once per geometry
# before the solver starts; the reference computes both on the host too.
# ----------------------------------------------------------------------------
def build_phi_from_stl_mm(mesh, nx: int, ny: int, nz: int, pitch_mm: float, origin_domain_mm, device="cuda",
                          dtype=torch.float32, chunk_size: int = 250_000):
    """Signed distance of the grid nodes origin + pitch * (i, j, k) to ``mesh``
    (trimesh's ProximityQuery.signed_distance, in chunks of ``chunk_size`` nodes in
    C order), as an (nx, ny, nz) tensor. Needs trimesh, as the reference does."""
    try:
        import trimesh  # noqa: F401
        from trimesh.proximity import ProximityQuery
    except ImportError as e:
        raise ImportError("build_phi_from_stl_mm needs trimesh (not installed)") from e
    pq = ProximityQuery(mesh)
    total = nx * ny * nz
    phi = np.empty(total, dtype=np.float32)
    org = np.asarray(origin_domain_mm, np.float64).reshape(3)
    for start in range(0, total, chunk_size):
        ids = np.arange(start, min(start + chunk_size, total))
        ijk = np.stack([ids // (ny * nz), (ids // nz) % ny, ids % nz], axis=1).astype(np.float64)
        phi[start:start + ids.size] = np.asarray(pq.signed_distance(org[None, :] + pitch_mm * ijk), np.float32)
    return torch.tensor(phi.reshape(nx, ny, nz), device=device, dtype=dtype)


def build_phi_from_voxel_mask_mm(voxel_solid: np.ndarray, pitch_mm: float, device="cuda", dtype=torch.float32):
    """Signed distance of a voxel mask (scipy's exact Euclidean distance transform):
    in fluid the distance to the nearest solid voxel, in solid minus the distance to
    the nearest fluid voxel, times the pitch."""
    from scipy.ndimage import distance_transform_edt
    solid = np.asarray(voxel_solid).astype(bool)
    phi = distance_transform_edt(~solid) * pitch_mm
    phi[solid] = -(distance_transform_edt(solid) * pitch_mm)[solid]
    return torch.tensor(phi, device=device, dtype=dtype)
