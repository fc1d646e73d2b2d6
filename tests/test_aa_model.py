"""CPU model of the fused AA step's data flow, driven by the same tables the
HIP kernel reads (fluxcore3d_amd.solver.aa_engine.build_tables), checked
against the FP64 oracle's reference step. This pins the table design (slot
addressing, ghost copies, solid-source substitution, Bouzidi A/B buffers) on
the CPU, independently of the GPU.
"""
import numpy as np
import pytest

from fluxcore3d_amd.solver.aa_engine import build_tables
from oracle import lbm as olbm
from tests import cases

C, OPP = olbm.C_NP, olbm.OPP


def _wrap_or_pad(c, d, n, g):
    v = c + d
    if g:
        return v + 1
    return np.where(v < 0, v + n, np.where(v >= n, v - n, v))


def _popc_below(mask, q):
    m = (mask & np.uint32((1 << q) - 1)).astype(np.uint64)
    cnt = np.zeros(m.shape, np.int64)
    for b in range(19):
        cnt += ((m >> np.uint64(b)) & np.uint64(1)).astype(np.int64)
    return cnt


def model_run(nx, ny, nz, solid, links, periodic, f0, steps, om, ibb_mode="reference"):
    tb = build_tables(nx, ny, nz, solid, links, periodic, ibb_mode)
    N = nx * ny * nz
    gx, gy, gz = (0 if p else 1 for p in periodic)
    PY, PZ = ny + 2 * gy, nz + 2 * gz
    Ns = (nx + 2 * gx) * PY * PZ
    ii, jj, kk = np.unravel_index(np.arange(N), (nx, ny, nz))

    def slot(d):
        return ((_wrap_or_pad(ii, d[0], nx, gx) * PY + _wrap_or_pad(jj, d[1], ny, gy)) * PZ
                + _wrap_or_pad(kk, d[2], nz, gz))

    def oob(q):
        src = (ii - C[q, 0], jj - C[q, 1], kk - C[q, 2])
        o = np.zeros(N, bool)
        for a, (g, n) in enumerate(((gx, nx), (gy, ny), (gz, nz))):
            if g:
                o |= (src[a] < 0) | (src[a] >= n)
        return o

    self_slot = slot((0, 0, 0))
    pop = np.zeros((19, Ns))
    pop[:, self_slot] = f0
    A = f0[tb["link_dir"], tb["link_node"]].astype(np.float64)
    B = [tb["B_init"].astype(np.float64), tb["B_init"].astype(np.float64)]
    fluid = ~solid.reshape(-1)
    sp = tb["sp_idx"]
    L, S, V, F = (tb["sp_mask"][:, a].view(np.uint32) for a in range(4))
    lb, vb, fb = tb["sp_base"][:, 0], tb["sp_base"][:, 1], tb["sp_base"][:, 2]
    w1, w2 = tb["w1"].astype(np.float64), tb["w2"].astype(np.float64)
    macros = None
    for t in range(steps):
        P, first = t & 1, t == 0
        f = np.empty((19, N))
        for q in range(19):
            f[q] = pop[q, self_slot] if P == 0 else pop[OPP[q], slot(-C[q])]
        for q in range(1, 19):
            hl = ((L >> np.uint32(q)) & np.uint32(1)).astype(bool)
            l = lb[hl] + _popc_below(L[hl], q)
            val = A[l].copy()
            if not first:
                hv = ((V[hl] >> np.uint32(q)) & np.uint32(1)).astype(bool)
                mv = vb[hl][hv] + _popc_below(V[hl][hv], q)
                val[hv] = val[hv] + w2[l[hv]] * B[P ^ 1][mv]
            f[q, sp[hl]] = val
            if not first:
                hs = ((S >> np.uint32(q)) & np.uint32(1)).astype(bool) & ~hl
                f[q, sp[hs]] = olbm.W32[q]
        fp, rho, u, v, w = olbm.macro_collide("bgk", f, (~fluid).astype(np.uint8), om1=om)
        macros = (rho, u, v, w)
        for q in range(19):
            if P == 0:
                pop[OPP[q], self_slot[fluid]] = fp[q, fluid]
                o = oob(q) & fluid
                pop[OPP[q], slot(-C[q])[o]] = fp[q, o]
            else:
                pop[q, slot(C[q])[fluid]] = fp[q, fluid]
                o = oob(q) & fluid
                pop[q, self_slot[o]] = fp[q, o]
        for q in range(1, 19):
            hl = ((L >> np.uint32(q)) & np.uint32(1)).astype(bool)
            l = lb[hl] + _popc_below(L[hl], q)
            x = sp[hl]
            hv = ((V[hl] >> np.uint32(q)) & np.uint32(1)).astype(bool)
            A[l] = w1[l] * fp[OPP[q], x] + np.where(hv, 0.0, w2[l] * fp[q, x])
            hf = ((F >> np.uint32(q)) & np.uint32(1)).astype(bool)
            e = fb[hf] + _popc_below(F[hf], q)
            B[P][tb["ff_idx"][e]] = fp[q, sp[hf]]
    return macros


@pytest.mark.parametrize("case", ["periodic_pack", "walled_box"])
def test_aa_model_matches_oracle(case):
    if case == "periodic_pack":
        c = cases.periodic_pack(n=24, seed=3)
        per = (1, 1, 1)
        solid = c["solid"]
    else:
        c = cases.config1()
        per = (0, 0, 0)
        solid = c["solid_flow"]
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    N = nx * ny * nz
    rng = np.random.default_rng(0)
    rho = 1 + 0.01 * rng.random(N)
    u = 0.02 * rng.standard_normal((3, N))
    f0 = olbm.feq(rho, *u, np.float64)
    om = 1.0 / 0.55
    steps = 6
    mm = model_run(nx, ny, nz, solid, c["links"], per, f0, steps, om)
    f = f0.copy()
    for _ in range(steps):
        f, *mo = olbm.step(f, solid, nx, ny, nz, per, "bgk", om, 1.0, None, 0.0, c["links"])
    for a, b in zip(mm, mo):
        assert np.abs(a - b).max() < 1e-12


def test_wave_offsets_match_per_node_bases():
    """The fused step derives each special node's record offsets from the per-group
    table (wave_meta) and an exclusive scan of record counts over the group's special
    nodes; on a porous pack this must equal the per-node bases (sp_base)."""
    c = cases.periodic_pack(n=24, seed=3)
    n = c["nx"]
    tb = build_tables(n, n, n, c["solid"], c["links"], (1, 1, 1))
    wm, sp, mask, base = tb["wave_meta"], tb["sp_idx"], tb["sp_mask"].view(np.uint32), tb["sp_base"]
    groups = (n ** 3 + 63) // 64
    assert wm.shape == (groups + 1, 4)
    assert tuple(wm[-1]) == (tb["n_special"], tb["n_links"], tb["n_v"], tb["ff_node"].size)
    pop = np.vectorize(lambda v: bin(int(v)).count("1"))
    for g in range(groups):
        sel = slice(wm[g, 0], wm[g + 1, 0])
        assert np.all(sp[sel] // 64 == g)
        m = mask[sel]
        for col, (a, k) in enumerate(((1, 0), (2, 2), (3, 3))):
            cnt = pop(m[:, k]) if m.size else np.zeros(0, int)
            offs = wm[g, a] + np.concatenate([[0], np.cumsum(cnt)[:-1]]) if cnt.size else cnt
            assert np.array_equal(offs, base[sel, col]), (g, col)
