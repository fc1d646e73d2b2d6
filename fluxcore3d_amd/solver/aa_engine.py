"""Fused in-place D3Q19 step (AA pattern) — host side.

``AAEngine`` owns the population buffer of the fused kernel in
``csrc/lbm_aa.hip`` and the small node/link tables it reads. It is built from
the same inputs as the reference step (lbm_cuda_ext.py:445-517): the flow
solid mask, the Bouzidi link set (``IBBData``, sdf_tracing.py:8-18), the
streaming periodicity and the list of inlet/outlet planes. One ``step()`` is
one launch of the fused kernel plus the plane kernels of the flow BCs; the
state it leaves equals the reference's ``sim.f`` after ``_step_cuda``.

Tables (built once, in node order):
  nflag[N]       uint8: 1 solid, 2 special, bits 3..5 first BC plane + 1
  wave_base[G]   special nodes before each group of 64 nodes
  sp_mask[S,4]   per special node: link mask L, solid-source mask S,
                 second-node mask V, publish mask F
  sp_base[S,4]   first link without a second node, first V-link slot in B,
                 first publish entry, first link with a second node
  w1, w2, A      per link: by wavefront (64 lanes), links with a second node
                 first, then by node and direction
  B[2, n_v]      second-node values, double-buffered by step parity
  ff_idx[n_ff]   V-link slot each publishing node writes
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from .. import _lib
from .d3q19 import C as C_NP, OPP as OPP_NP, W as W_NP

STORAGE_CODES = {"fp32": 0, "fp16": 1, "bf16": 2}
STORAGE_DTYPES = {"fp32": torch.float32, "fp16": torch.float16, "bf16": torch.bfloat16}
COLLISION_CODES = {"bgk": 0, "trt": 1, "mrt": 2, "mrt_smag": 3}
# Compact storage: the stride between the 19 direction arrays is the slot count rounded
# up to this many slots, so that every direction array starts on a 256-byte boundary
# (FP32) and a wavefront's 64 consecutive own slots fill whole 128-byte lines.
SLOT_ALIGN = 64

F_SOLID, F_SPECIAL, F_BC_SHIFT = 1, 2, 3


def _neighbor_solid_masks(solid: np.ndarray, periodic) -> np.ndarray:
    """bit q set where the upstream node x - c_q is solid (wrapped on periodic
    axes, out-of-domain counts as fluid)."""
    nx, ny, nz = solid.shape
    S = np.zeros(solid.shape, dtype=np.uint32)
    pad = np.pad(solid, 1, mode="constant", constant_values=False)
    for ax, per in enumerate(periodic):
        if per:
            # wrap the halo on periodic axes
            idx = [slice(None)] * 3
            idx[ax] = 0
            src = [slice(None)] * 3
            src[ax] = -2
            pad[tuple(idx)] = pad[tuple(src)]
            idx[ax] = -1
            src[ax] = 1
            pad[tuple(idx)] = pad[tuple(src)]
    for q in range(1, 19):
        cx, cy, cz = C_NP[q]
        up = pad[1 - cx:1 - cx + nx, 1 - cy:1 - cy + ny, 1 - cz:1 - cz + nz]
        S |= up.astype(np.uint32) << np.uint32(q)
    return S


def build_storage(nx, ny, nz, solid: np.ndarray, periodic, bcs=()) -> dict:
    """Compact population storage: which nodes of the padded grid own slots, and
    the per-row records that turn padded coordinates into storage indices.

    A node owns slots (is present) when it is a fluid node, a ghost node of a
    non-periodic face (copy-self values) or a solid node upstream of a node on an
    inlet/outlet plane (the plane kernels rebuild every input of their nodes,
    including those streamed from solids). Inputs streamed from a node without slots
    are w_q (the reference streams f_post = W out of solids, lbm_kernels.cu:277-283).

    Storage index of padded node (I, J, K): with R = I * PY + J and word w = K / 32,
    row_meta[R * wpr + w] = (mask, start): bit b of mask is the presence of K = 32 w + b
    and start counts the present nodes before the word, so the index is
    start + popcount(mask & ((1 << b) - 1)). Nodes are numbered in padded C order."""
    solid = np.asarray(solid, bool).reshape(nx, ny, nz)
    g = [0 if p else 1 for p in periodic]
    PX, PY, PZ = nx + 2 * g[0], ny + 2 * g[1], nz + 2 * g[2]
    present = np.ones((PX, PY, PZ), bool)
    present[g[0]:g[0] + nx, g[1]:g[1] + ny, g[2]:g[2] + nz] = ~solid
    plane_nodes = np.zeros((nx, ny, nz), bool)
    for bc in bcs:
        sel = [slice(None)] * 3
        sel[int(bc["axis"])] = int(bc["plane"])
        plane_nodes[tuple(sel)] = True
    plane_nodes &= ~solid
    if plane_nodes.any():
        ii, jj, kk = np.nonzero(plane_nodes)
        for q in range(1, 19):
            c = C_NP[q]
            src = []
            for a, (v, n) in enumerate(((ii, nx), (jj, ny), (kk, nz))):
                w = v - c[a]
                src.append(w + 1 if g[a] else np.mod(w, n))
            present[src[0], src[1], src[2]] = True
    wpr = (PZ + 31) // 32
    bits = np.zeros((PX * PY, wpr * 32), bool)
    bits[:, :PZ] = present.reshape(PX * PY, PZ)
    words = bits.reshape(-1, 32)
    mask = (words.astype(np.uint64) << np.arange(32, dtype=np.uint64)).sum(1).astype(np.uint32)
    cnt = words.sum(1).astype(np.int64)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    n_store = int(cnt.sum())
    if n_store >= 2 ** 30:
        raise ValueError("compact storage of more than 2^30 nodes")
    row_meta = np.stack([mask.view(np.int32), start.astype(np.int32)], 1)
    return dict(row_meta=row_meta, wpr=wpr, n_store=n_store, present=present, PY=PY, PZ=PZ, gpad=tuple(g))


def build_tables(nx, ny, nz, solid: np.ndarray, links, periodic, ibb_mode="reference", bcs=(),
                 compact: bool = True, lane_mask: np.ndarray | None = None) -> dict:
    """Node/link tables of the fused step (pure NumPy, node order).

    Lanes: one per node (``compact=False``), one per fluid node (``compact=True``,
    dense storage: inputs from solid nodes are substituted from the source list), or
    one per node of ``lane_mask`` (compact storage: the nodes that own slots; an
    input from a node without slots reads as w_q, so there is no source list)."""
    N = nx * ny * nz
    solid = np.asarray(solid, bool).reshape(nx, ny, nz)
    flat_solid = solid.reshape(-1)
    Smask = _neighbor_solid_masks(solid, periodic).reshape(-1)
    Smask[flat_solid] = 0
    Lmask = np.zeros(N, dtype=np.uint32)
    Vmask = np.zeros(N, dtype=np.uint32)
    Fmask = np.zeros(N, dtype=np.uint32)
    w1s = np.zeros(0, np.float32)
    w2s = np.zeros(0, np.float32)
    link_node = np.zeros(0, np.int64)
    link_dir = np.zeros(0, np.int64)
    v_node = np.zeros(0, np.int64)
    ff_node = np.zeros(0, np.int64)
    ff_dir = np.zeros(0, np.int64)
    ff_slot = np.zeros(0, np.int32)
    link_v = np.zeros(0, bool)
    B_init = np.zeros(0, np.float32)
    if links is not None and int(np.asarray(links["fluid_flat"]).size) > 0:
        xf = np.asarray(links["fluid_flat"], np.int64)
        q = np.asarray(links["dir_i"], np.int64)
        w1 = np.asarray(links["w1"], np.float32).copy()
        w2 = np.asarray(links["w2"], np.float32).copy()
        ge = np.asarray(links["q_ge_half"], bool)
        ffv = np.asarray(links["ff_valid"], bool)
        ff = np.asarray(links["ff_flat"], np.int64)
        if xf.min() < 0 or xf.max() >= N or q.min() < 1 or q.max() > 18:
            raise ValueError("link table out of range")
        keep = ~flat_solid[xf]  # links at flow-solid nodes are reset by the reference
        xf, q, w1, w2, ge, ffv, ff = xf[keep], q[keep], w1[keep], w2[keep], ge[keep], ffv[keep], ff[keep]
        t = q if ibb_mode == "reference" else OPP_NP[q]
        bad = (~ge) & (~ffv)  # q < 1/2 without a second node: f*_opp alone
        w1[bad] = 1.0
        w2[bad] = 0.0
        v = (~ge) & ffv
        order = np.lexsort((t, xf))
        xf, t, w1, w2, v, ff = xf[order], t[order], w1[order], w2[order], v[order], ff[order]
        key = xf * 19 + t
        if np.any(key[1:] == key[:-1]):
            raise ValueError("duplicate Bouzidi link (node, direction)")
        np.bitwise_or.at(Lmask, xf, (np.uint32(1) << t.astype(np.uint32)))
        Smask[xf] &= ~Lmask[xf]  # a link overrides a solid-source substitution
        np.bitwise_or.at(Vmask, xf[v], (np.uint32(1) << t[v].astype(np.uint32)))
        w1s, w2s, link_node, link_dir = w1, w2, xf, t
        vx, vt, vff = xf[v], t[v], ff[v]
        v_node = vx
        n_v = int(vx.size)
        ff_solid = flat_solid[vff]
        p = OPP_NP[vt]  # population the second node publishes
        B_init = np.where(ff_solid, W_NP[p], 0.0).astype(np.float32)
        pub = ~ff_solid
        pm = np.arange(n_v, dtype=np.int64)[pub]
        pn, pp = vff[pub], p[pub]
        o2 = np.lexsort((pp, pn))
        pn, pp, pm = pn[o2], pp[o2], pm[o2]
        if np.any((pn * 19 + pp)[1:] == (pn * 19 + pp)[:-1]):
            raise ValueError("duplicate second-node publication")
        np.bitwise_or.at(Fmask, pn, (np.uint32(1) << pp.astype(np.uint32)))
        ff_node, ff_dir, ff_slot = pn, pp, pm.astype(np.int32)
        link_v = v
    special = (~flat_solid) & ((Lmask | Smask | Fmask) != 0)
    sp_idx = np.nonzero(special)[0]

    def first_index(owner: np.ndarray) -> np.ndarray:
        return np.searchsorted(owner, sp_idx).astype(np.int32)

    sp_mask = np.stack([Lmask[sp_idx], Smask[sp_idx], Vmask[sp_idx], Fmask[sp_idx]], 1).view(np.int32)
    sp_base = np.stack([np.zeros(sp_idx.size, np.int32), first_index(v_node), first_index(ff_node),
                        np.zeros(sp_idx.size, np.int32)], 1).astype(np.int32)
    nflag = flat_solid.astype(np.uint8) * F_SOLID | special.astype(np.uint8) * F_SPECIAL
    for b, bc in enumerate(bcs):
        ax, pl = int(bc["axis"]), int(bc["plane"])
        sel = [slice(None)] * 3
        sel[ax] = pl
        plane = np.zeros((nx, ny, nz), bool)
        plane[tuple(sel)] = True
        plane = plane.reshape(-1) & ~flat_solid & ((nflag >> F_BC_SHIFT) == 0)
        nflag[plane] |= np.uint8((b + 1) << F_BC_SHIFT)
    groups = (N + 63) // 64
    cnt = np.zeros(groups * 64, np.int64)
    cnt[:N] = special
    wave_base = np.concatenate([[0], np.cumsum(cnt.reshape(groups, 64).sum(1))[:-1]]).astype(np.int32)
    # Lanes of the fused step: with `compact`, one lane per fluid node (fluid nodes in
    # node order, 64 per wavefront, `lane_nodes`); otherwise one lane per node (solid
    # lanes load whole lines and store w_q). Entry lists are node-major, so each wavefront owns a contiguous
    # range of every list. Per-entry codes:
    #   link:    lane | t << 6 | opp(t) << 11 | second node << 16
    #            | override << 17 | second-node index relative to the wavefront's first << 18
    #   source:  lane | q << 6 | override << 17   (solid upstream, input w_q after the first step;
    #            compacted lanes only: with one lane per node, solid lanes write w_q into the
    #            slots their neighbours read, so the list is empty)
    #   publish: lane | p << 6                     (post-collision f_p goes to B[ff_idx])
    # override = the node is not on an inlet/outlet plane (plane kernels rebuild those inputs).
    if lane_mask is not None:
        lane_mask = np.asarray(lane_mask, bool).reshape(-1)
        if np.any(~lane_mask & ~flat_solid):
            raise ValueError("every fluid node needs a lane")
        lane_nodes = np.nonzero(lane_mask)[0].astype(np.int64)
    elif compact:
        lane_nodes = np.nonzero(~flat_solid)[0].astype(np.int64)
    else:
        lane_nodes = np.arange(N, dtype=np.int64)
    n_lanes = int(lane_nodes.size)
    rank = np.full(N, -1, np.int64)
    rank[lane_nodes] = np.arange(n_lanes)
    waves = (n_lanes + 63) // 64
    starts = np.append(lane_nodes[::64], N) if n_lanes else np.zeros(1, np.int64)
    ovr_node = (nflag >> F_BC_SHIFT) == 0
    # Within each wavefront the links with a second node come first, so entry e < nV
    # of a wavefront is its e-th second-node link (B slot v_lo + e) and the kernel
    # loads B without waiting for the entry code. The second-node links keep the
    # global (node, direction) order of the B slots.
    link_wave = rank[link_node] // 64
    lo = np.lexsort((link_dir, link_node, ~link_v, link_wave))
    link_node, link_dir, w1s, w2s, link_v, link_wave = (a[lo] for a in (link_node, link_dir, w1s, w2s, link_v,
                                                                        link_wave))
    assert np.array_equal(link_node[link_v], v_node)
    pos = np.arange(link_node.size)

    def first_pos(sel: np.ndarray) -> np.ndarray:
        # first position of each special node's links in the class `sel` (node-sorted)
        return np.append(pos[sel], link_node.size)[np.searchsorted(link_node[sel], sp_idx)].astype(np.int32)

    sp_base[:, 0] = first_pos(~link_v)
    sp_base[:, 3] = first_pos(link_v)
    v_base = np.searchsorted(v_node, starts)
    v_rank = np.cumsum(link_v) - 1
    lk_code = ((rank[link_node] % 64) | (link_dir << 6) | (OPP_NP[link_dir] << 11)
               | (link_v.astype(np.int64) << 16) | (ovr_node[link_node].astype(np.int64) << 17)
               | (np.where(link_v, v_rank - v_base[rank[link_node] // 64], 0) << 18)).astype(np.int32)
    if compact and lane_mask is None:
        s_row, s_q = np.nonzero(((Smask[sp_idx][:, None] >> np.arange(19, dtype=np.uint32)) & 1).astype(bool))
        s_node = sp_idx[s_row]
    else:
        s_q = s_node = np.zeros(0, np.int64)
    s_code = ((rank[s_node] % 64) | (s_q << 6) | (ovr_node[s_node].astype(np.int64) << 17)).astype(np.int32)
    f_code = ((rank[ff_node] % 64) | (ff_dir << 6)).astype(np.int64)
    # per wavefront (+ totals): first link, second-node link, publish entry and source entry
    wave_meta = np.stack([np.searchsorted(link_wave, np.arange(waves + 1)), v_base,
                          np.searchsorted(ff_node, starts), np.searchsorted(s_node, starts)], 1).astype(np.int32)
    assert wave_meta.shape[0] == waves + 1 or n_lanes == 0
    # A publication's B slot relative to its group's first second-node slot fits in the
    # code's upper 21 bits on any grid whose x-planes hold under about a million links:
    # the publishing node is a neighbour of the consuming one. Otherwise ff_idx is kept.
    rel = ff_slot.astype(np.int64) - wave_meta[rank[ff_node] // 64, 1] if ff_node.size else np.zeros(0, np.int64)
    ff_packed = bool(rel.size == 0 or (rel.min() >= -(1 << 20) and rel.max() < (1 << 20)))
    if ff_packed:
        f_code = f_code | (rel << 11)
    f_code = f_code.astype(np.int32)  # in range: lane | p << 6 < 2^11, rel << 11 within 32 bits
    assert np.all((pos - wave_meta[link_wave, 0] < np.diff(wave_meta[:, 1])[link_wave]) == link_v)
    return dict(nflag=nflag, wave_base=wave_base, wave_meta=wave_meta, lk_code=lk_code, s_code=s_code,
                f_code=f_code, ff_packed=ff_packed, n_src=int(s_node.size), lane_nodes=lane_nodes.astype(np.int32), n_lanes=n_lanes, sp_mask=sp_mask, sp_base=sp_base, w1=w1s, w2=w2s,
                link_node=link_node, link_dir=link_dir, B_init=B_init, ff_idx=ff_slot, ff_node=ff_node,
                n_links=int(link_node.size), n_v=int(v_node.size), n_special=int(sp_idx.size), sp_idx=sp_idx)


class AAEngine:
    """In-place D3Q19 state + tables of the fused kernel."""

    def __init__(self, nx, ny, nz, solid_flow, *, periodic=(False, False, False), collision="bgk",
                 omega=1.0, omega_minus=1.0, S=None, cs_dx2=0.0, force=None, sponge=None, sponge_axis=0,
                 links=None, ibb_mode="reference", bcs=(), storage="fp32", device="cuda", lanes="auto",
                 layout="auto", compress_links=True):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.N = self.nx * self.ny * self.nz
        self.periodic = tuple(bool(p) for p in periodic)
        if storage not in STORAGE_CODES:
            raise ValueError(f"storage must be one of {sorted(STORAGE_CODES)}")
        if collision not in COLLISION_CODES:
            raise ValueError(f"collision must be one of {sorted(COLLISION_CODES)}")
        if ibb_mode not in ("reference", "textbook"):
            raise ValueError("ibb_mode must be 'reference' or 'textbook'")
        self.storage = storage
        solid = np.ascontiguousarray(np.asarray(solid_flow, dtype=bool).reshape(self.nx, self.ny, self.nz))
        # Population layout: "dense" (every node of the padded grid owns 19 slots) or
        # "compact" (only fluid, ghost and plane-feeding solid nodes own slots,
        # build_storage). "auto": compact for FP32 populations when at least 5 % of
        # the nodes are solid (256^3 sphere pack, 40 % solid: 0.565 against 0.626 ms
        # per step; 192x192x384 heat sink, 12.5 % solid: 0.718 against 0.848 ms),
        # dense for 16-bit populations (the pack: 0.519 dense against 0.732 ms compact,
        # half the bytes per lane leave too few in flight per wavefront).
        if layout == "auto":
            layout = "compact" if storage == "fp32" and solid.mean() >= 0.05 else "dense"
        if layout not in ("dense", "compact"):
            raise ValueError("layout must be 'auto', 'dense' or 'compact'")
        self.layout = layout
        # compress_links: the kernel forms w2 = 1 - w1 where that holds bit for bit and
        # reads each publication's B slot from its code where it fits (4 + 4 B less per
        # link and publication); False keeps the explicit w2 and ff_idx arrays.
        self._compress_links = bool(compress_links)
        # Lane mapping of the fused step: "fluid" or "all" (one lane per node).
        # Dense storage: "fluid" is one lane per fluid node; "all" (the default) is
        # 1.8x (FP32) and 1.5x (FP16) faster on a porosity-0.6 256^3 sphere pack,
        # because each load instruction covers whole cache lines instead of gathering
        # scattered nodes. Compact storage: "fluid" (the default) is one lane per node
        # that owns slots (fluid nodes and the solid nodes feeding inlet/outlet planes).
        # These lanes read consecutive slots, and every lane of a wavefront has
        # populations to move, so each wavefront keeps 64 lanes x 19 loads in flight;
        # with "all" a third of the lanes idle on solid nodes.
        if lanes == "auto":
            lanes = "fluid" if layout == "compact" else "all"
        if lanes not in ("fluid", "all"):
            raise ValueError("lanes must be 'auto', 'fluid' or 'all'")
        self.lanes = lanes
        self.device = torch.device(device)
        self.ibb_mode = ibb_mode
        self.bcs = [dict(b) for b in bcs]
        if len(self.bcs) > 6:
            raise ValueError("at most 6 inlet/outlet planes")
        self._build_tables(solid, links)
        self._collision = self._make_collision(collision, omega, omega_minus, S, cs_dx2, force, sponge, sponge_axis)
        if self.layout == "compact":
            self.Ns = self._n_store
        else:
            self.Ns = int(_lib.call("fc_aa_slot_count", self.nx, self.ny, self.nz,
                                    *[int(p) for p in self.periodic]))
        self.pop = torch.zeros(19 * self.Ns, dtype=STORAGE_DTYPES[storage], device=self.device)
        self.t = 0
        self.t_loaded = 0
        self._ws = None
        self._macro_ptrs = None

    # ------------------------------------------------------------------ tables
    def _build_tables(self, solid: np.ndarray, links) -> None:
        sb = None
        lane_mask = None
        if self.layout == "compact":
            sb = build_storage(self.nx, self.ny, self.nz, solid, self.periodic, self.bcs)
            if self.lanes == "fluid":
                g = sb["gpad"]
                lane_mask = sb["present"][g[0]:g[0] + self.nx, g[1]:g[1] + self.ny, g[2]:g[2] + self.nz]
        tb = build_tables(self.nx, self.ny, self.nz, solid, links, self.periodic, self.ibb_mode, self.bcs,
                          compact=self.lanes == "fluid", lane_mask=lane_mask)
        self.n_links, self.n_v, self.n_special = tb["n_links"], tb["n_v"], tb["n_special"]
        if sb is not None:
            stride = -(-sb["n_store"] // SLOT_ALIGN) * SLOT_ALIGN
            dense = int(_lib.call("fc_aa_slot_count", self.nx, self.ny, self.nz, *[int(p) for p in self.periodic]))
            if stride > dense:  # tiny grids: no room to round up
                stride = sb["n_store"]
            tb.update(row_meta=sb["row_meta"], wpr=sb["wpr"], n_store=stride)
            self._n_store = stride
        else:
            tb.update(row_meta=None, wpr=0, n_store=0)
        self._upload(tb)

    def _upload(self, tb) -> None:
        dev = self.device

        def dt(a, dtype):
            a = np.ascontiguousarray(a)
            if a.size == 0:
                a = np.zeros(1, a.dtype)
            return torch.from_numpy(a).to(device=dev, dtype=dtype)

        self.nflag = dt(tb["nflag"], torch.uint8)
        self.wave_base = dt(tb["wave_base"], torch.int32)
        self.wave_meta = dt(tb["wave_meta"], torch.int32)
        self.n_lanes = int(tb["n_lanes"])
        # the identity lane list (one lane per node) is not stored: lane r is node r
        identity = self.n_lanes == self.N
        self.lane_nodes = None if identity else dt(tb["lane_nodes"], torch.int32)
        self.lk_code = dt(tb["lk_code"], torch.int32)
        self.s_code = dt(tb["s_code"], torch.int32)
        self.f_code = dt(tb["f_code"], torch.int32)
        self.sp_mask = dt(tb["sp_mask"], torch.int32)
        self.sp_base = dt(tb["sp_base"], torch.int32)
        self.w1 = dt(tb["w1"], torch.float32)
        self.w2 = dt(tb["w2"], torch.float32)
        self.A = torch.zeros(max(self.n_links, 1), dtype=torch.float32, device=dev)
        B_init = tb["B_init"]
        self.B_init = dt(np.concatenate([B_init, B_init]) if self.n_v else np.zeros(2, np.float32), torch.float32)
        self.B = self.B_init.clone()
        self.ff_idx = dt(tb["ff_idx"], torch.int32)
        # tables the kernel can do without: w2 where it is 1 - w1 bit for bit (the
        # reference's weights), ff_idx where the publication codes carry the slot
        w1, w2 = np.asarray(tb["w1"], np.float32), np.asarray(tb["w2"], np.float32)
        self.w2_derived = self._compress_links and bool(np.array_equal(w2, np.float32(1.0) - w1))
        self.ff_packed = self._compress_links and bool(tb["ff_packed"])
        self.row_meta = None if tb["row_meta"] is None else dt(tb["row_meta"], torch.int32)
        self._tables = _lib.AATables(
            self.nflag.data_ptr(), self.wave_base.data_ptr(), self.wave_meta.data_ptr(), self.lk_code.data_ptr(),
            self.s_code.data_ptr(), self.f_code.data_ptr(), _lib.ptr(self.lane_nodes), self.n_lanes,
            self.sp_mask.data_ptr(), self.sp_base.data_ptr(),
            self.w1.data_ptr(), 0 if self.w2_derived else self.w2.data_ptr(), self.A.data_ptr(), self.B.data_ptr(),
            0 if self.ff_packed else self.ff_idx.data_ptr(),
            max(self.n_v, 1), _lib.ptr(self.row_meta), int(tb["wpr"]), int(tb["n_store"]))

    def _make_collision(self, collision, omega, omega_minus, S, cs_dx2, force, sponge, sponge_axis):
        c = _lib.AACollision()
        c.kind = COLLISION_CODES[collision]
        c.om1 = float(omega)
        c.om2 = float(omega_minus)
        if c.kind >= 2:
            if S is None:
                raise ValueError("MRT collision needs the 19 relaxation rates S")
            s = np.asarray(S, np.float32).reshape(-1)
            if s.size != 19:
                raise ValueError("S must have 19 elements")
            for k in range(19):
                c.S[k] = float(s[k])
        c.cs_dx2 = float(cs_dx2)
        if force is not None and any(float(v) != 0.0 for v in force):
            for d in range(3):
                c.force[d] = float(force[d])
            c.use_force = 1
        self._sponge = None
        if sponge is not None:
            prof = torch.as_tensor(np.asarray(sponge, np.float32)).to(self.device)
            if prof.numel() != (self.nx, self.ny, self.nz)[sponge_axis]:
                raise ValueError("sponge profile length must match the sponge axis")
            self._sponge = prof.contiguous()
            c.sponge = self._sponge.data_ptr()
            c.sponge_axis = int(sponge_axis)
        return c

    # ------------------------------------------------------------------ state
    @property
    def _dims(self):
        return (self.nx, self.ny, self.nz, *[int(p) for p in self.periodic])

    def load(self, f: torch.Tensor) -> None:
        """Start from reference-layout populations f[19, nx, ny, nz] (FP32)."""
        if f.dtype != torch.float32 or f.numel() != 19 * self.N:
            raise RuntimeError("f must be float32 with 19*nx*ny*nz elements")
        f = f.to(self.device).contiguous()
        self.B.copy_(self.B_init)
        _lib.call("fc_aa_load", self.pop.data_ptr(), STORAGE_CODES[self.storage], *self._dims, self._tables,
                  f.data_ptr(), _lib.stream_handle(self.device))
        self.t_loaded = self.t

    def materialize(self, out: torch.Tensor | None = None) -> torch.Tensor:
        """Populations entering the next step, reference layout [19, nx, ny, nz]."""
        if out is None:
            out = torch.empty((19, self.nx, self.ny, self.nz), dtype=torch.float32, device=self.device)
        _lib.call("fc_aa_materialize", self.pop.data_ptr(), STORAGE_CODES[self.storage], *self._dims,
                  self._tables, self.t & 1, int(self.t == self.t_loaded), out.data_ptr(),
                  _lib.stream_handle(self.device))
        return out

    def step(self, rho=None, u=None, v=None, w=None, solid_macros: Optional[bool] = None) -> None:
        """One time step; writes pre-collision macros if the four outputs are given.
        Solid nodes hold the constants (1, 0, 0, 0); they are written when the
        output buffers differ from the previous step's and skipped otherwise, or as
        ``solid_macros`` says (a caller that alternates between buffer sets whose solid
        entries it has already written passes False)."""
        self.step_fused(rho, u, v, w, solid_macros)
        self.apply_bcs()

    def step_fused(self, rho=None, u=None, v=None, w=None, solid_macros: Optional[bool] = None) -> None:
        """The fused collide/stream/Bouzidi launch of step(), without the inlet/outlet
        plane kernels (apply_bcs() must follow before the next step)."""
        stream = _lib.stream_handle(self.device)
        ptrs = (_lib.ptr(rho), _lib.ptr(u), _lib.ptr(v), _lib.ptr(w))
        solid_macros = int(ptrs != self._macro_ptrs) if solid_macros is None else int(bool(solid_macros))
        self._macro_ptrs = ptrs
        _lib.call("fc_aa_step", self.pop.data_ptr(), STORAGE_CODES[self.storage], *self._dims, self._tables,
                  self._collision, self.t & 1, int(self.t == self.t_loaded), *ptrs, solid_macros, stream)
        self.t += 1

    def apply_bcs(self) -> None:
        """Inlet/outlet plane kernels on the inputs of the next step."""
        self._apply_bcs(_lib.stream_handle(self.device))

    def _apply_bcs(self, stream) -> None:
        parity = self.t & 1
        for b, bc in enumerate(self.bcs):
            if bc["kind"] == "zou_he":
                _lib.call("fc_aa_zou_he", self.pop.data_ptr(), STORAGE_CODES[self.storage], *self._dims,
                          self._tables, parity, 0, b, int(bc["axis"]), int(bc["plane"]), int(bc["interior"]),
                          int(bc["bc_type"]), float(bc["u_in"]), float(bc["rho_out"]), int(bc["sign"]), stream)
            elif bc["kind"] == "convective":
                n1, n2 = {0: (self.ny, self.nz), 1: (self.nx, self.nz), 2: (self.nx, self.ny)}[int(bc["axis"])]
                need = int(_lib.call("fc_aa_convective_workspace_bytes", n1, n2))
                if self._ws is None or self._ws.numel() < need:
                    self._ws = torch.empty(max(need, 16), dtype=torch.uint8, device=self.device)
                _lib.call("fc_aa_convective", self.pop.data_ptr(), STORAGE_CODES[self.storage], *self._dims,
                          self._tables, parity, 0, b, int(bc["axis"]), int(bc["plane"]), int(bc["interior"]),
                          int(bc["unknown_cax"]), float(bc["rho_out"]), self._ws.data_ptr(), self._ws.numel(),
                          stream)
            else:
                raise ValueError(f"unknown BC kind {bc['kind']!r}")

    # ------------------------------------------------------------------ info
    def bytes_per_update(self) -> int:
        """Algorithmic population bytes per lattice update (19 loads + 19 stores)."""
        return 2 * 19 * self.pop.element_size()
