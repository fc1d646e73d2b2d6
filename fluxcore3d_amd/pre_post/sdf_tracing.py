"""Bouzidi link set and wall distances on the GPU (drop-in for pre_post/sdf_tracing.py).

``build_links_from_mask_gpu`` and ``build_ibb_from_sdf_gpu`` keep the
reference's names, arguments and ``IBBData`` result (sdf_tracing.py:8-259);
the per-node classification and the q search run in ``csrc/links.hip``.
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
