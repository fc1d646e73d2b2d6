"""CPU model of the fused AA step's data flow, driven by the same tables the
HIP kernel reads (fluxcore3d_amd.solver.aa_engine.build_tables), checked
against the FP64 oracle's reference step. This pins the table design (slot
addressing, ghost copies, solid-source substitution, Bouzidi A/B buffers) on
the CPU, independently of the GPU.
"""
import numpy as np
import pytest

from fluxcore3d_amd.solver.aa_engine import build_storage, build_tables
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


def _group_of(entry_base, n_entries):
    """Group (64-node wavefront) of each entry from the per-group first-entry table."""
    return np.searchsorted(entry_base, np.arange(n_entries), side="right") - 1


def _locate(sb, I, J, K):
    """Presence and storage index of padded nodes, as the fused kernel computes them
    from the per-row records (mask, start) of build_storage."""
    rec = sb["row_meta"][(I * sb["PY"] + J) * sb["wpr"] + (K >> 5)]
    mask = rec[:, 0].view(np.uint32).astype(np.uint64)
    b = (K & 31).astype(np.uint64)
    below = mask & ((np.uint64(1) << b) - np.uint64(1))
    cnt = np.zeros(I.shape, np.int64)
    for bit in range(32):
        cnt += ((below >> np.uint64(bit)) & np.uint64(1)).astype(np.int64)
    return ((mask >> b) & np.uint64(1)).astype(bool), rec[:, 1].astype(np.int64) + cnt


def model_run(nx, ny, nz, solid, links, periodic, f0, steps, om, ibb_mode="reference", compact=True,
              storage="dense", slot_lanes=False):
    lane_mask = None
    if storage == "compact" and slot_lanes:  # one lane per node owning slots
        sb = build_storage(nx, ny, nz, solid, periodic)
        g = sb["gpad"]
        lane_mask = sb["present"][g[0]:g[0] + nx, g[1]:g[1] + ny, g[2]:g[2] + nz]
    tb = build_tables(nx, ny, nz, solid, links, periodic, ibb_mode, compact=compact, lane_mask=lane_mask)
    if storage == "compact":
        return _model_run_compact(nx, ny, nz, solid, links, periodic, f0, steps, om, tb)
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

    wm = tb["wave_meta"].astype(np.int64)
    lk, sc, fc = (tb[k].astype(np.int64) for k in ("lk_code", "s_code", "f_code"))
    # decode the entry codes exactly as the fused kernel does
    fl = tb["lane_nodes"].astype(np.int64)
    l_node = fl[_group_of(wm[:, 0], lk.size) * 64 + (lk & 63)]
    l_t, l_opp = (lk >> 6) & 31, (lk >> 11) & 31
    l_hasv, l_ovr = ((lk >> 16) & 1).astype(bool), ((lk >> 17) & 1).astype(bool)
    l_v = wm[_group_of(wm[:, 0], lk.size), 1] + (lk >> 18)
    s_node = fl[_group_of(wm[:, 3], sc.size) * 64 + (sc & 63)]
    s_q, s_ovr = (sc >> 6) & 31, ((sc >> 17) & 1).astype(bool)
    f_node = fl[_group_of(wm[:, 2], fc.size) * 64 + (fc & 63)]
    f_p = (fc >> 6) & 31

    self_slot = slot((0, 0, 0))
    pop = np.zeros((19, Ns))
    pop[:, self_slot] = f0
    A = f0[tb["link_dir"], tb["link_node"]].astype(np.float64)
    B = [tb["B_init"].astype(np.float64), tb["B_init"].astype(np.float64)]
    fluid = ~solid.reshape(-1)
    # with one lane per node, solid lanes store their f_post (= w_q) like fluid lanes
    writer = fluid if compact else np.ones(N, bool)
    w1, w2 = tb["w1"].astype(np.float64), tb["w2"].astype(np.float64)
    macros = None
    for t in range(steps):
        P, first = t & 1, t == 0
        f = np.empty((19, N))
        for q in range(19):
            f[q] = pop[q, self_slot] if P == 0 else pop[OPP[q], slot(-C[q])]
        val = A.copy()
        if not first:
            val[l_hasv] += w2[l_hasv] * B[P ^ 1][l_v[l_hasv]]
        f[l_t[l_ovr], l_node[l_ovr]] = val[l_ovr]
        if not first:
            f[s_q[s_ovr], s_node[s_ovr]] = olbm.W32[s_q[s_ovr]]
        fp, rho, u, v, w = olbm.macro_collide("bgk", f, (~fluid).astype(np.uint8), om1=om)
        macros = (rho, u, v, w)
        for q in range(19):
            if P == 0:
                pop[OPP[q], self_slot[writer]] = fp[q, writer]
                o = oob(q) & fluid
                pop[OPP[q], slot(-C[q])[o]] = fp[q, o]
            else:
                pop[q, slot(C[q])[writer]] = fp[q, writer]
                o = oob(q) & fluid
                pop[q, self_slot[o]] = fp[q, o]
        A = w1 * fp[l_opp, l_node] + np.where(l_hasv, 0.0, w2 * fp[l_t, l_node])
        B[P][tb["ff_idx"]] = fp[f_p, f_node]
    return macros


def _model_run_compact(nx, ny, nz, solid, links, periodic, f0, steps, om, tb):
    """Compact storage, one lane per node or per node owning slots (decoded through
    the lane list, as the kernel does): lanes whose node has no slots do nothing; an
    input whose writer x - c_q has no slots is w_q (after the first step); a store
    to x + c_q happens only where that node has slots."""
    sb = build_storage(nx, ny, nz, solid, periodic)
    N = nx * ny * nz
    gx, gy, gz = sb["gpad"]
    ii, jj, kk = np.unravel_index(np.arange(N), (nx, ny, nz))

    def loc(d):
        return _locate(sb, _wrap_or_pad(ii, d[0], nx, gx), _wrap_or_pad(jj, d[1], ny, gy),
                       _wrap_or_pad(kk, d[2], nz, gz))

    def oob(q):
        src = (ii - C[q, 0], jj - C[q, 1], kk - C[q, 2])
        o = np.zeros(N, bool)
        for a, (g, n) in enumerate(((gx, nx), (gy, ny), (gz, nz))):
            if g:
                o |= (src[a] < 0) | (src[a] >= n)
        return o

    wm = tb["wave_meta"].astype(np.int64)
    lk, fc = tb["lk_code"].astype(np.int64), tb["f_code"].astype(np.int64)
    assert tb["n_src"] == 0  # inputs from nodes without slots read as w_q: no source list
    fl = tb["lane_nodes"].astype(np.int64)
    l_node = fl[_group_of(wm[:, 0], lk.size) * 64 + (lk & 63)]
    l_t, l_opp = (lk >> 6) & 31, (lk >> 11) & 31
    l_hasv, l_ovr = ((lk >> 16) & 1).astype(bool), ((lk >> 17) & 1).astype(bool)
    l_v = wm[_group_of(wm[:, 0], lk.size), 1] + (lk >> 18)
    f_node = fl[_group_of(wm[:, 2], fc.size) * 64 + (fc & 63)]
    f_p = (fc >> 6) & 31
    self_p, self_i = loc((0, 0, 0))
    up = [loc(-C[q]) for q in range(19)]  # writer x - c_q of input q
    assert np.array_equal(self_p, ~solid.reshape(-1))
    pop = np.full((19, sb["n_store"]), np.nan)
    pop[:, self_i[self_p]] = f0[:, self_p]
    A = f0[tb["link_dir"], tb["link_node"]].astype(np.float64)
    B = [tb["B_init"].astype(np.float64), tb["B_init"].astype(np.float64)]
    fluid = ~solid.reshape(-1)
    w1, w2 = tb["w1"].astype(np.float64), tb["w2"].astype(np.float64)
    macros = None
    for t in range(steps):
        P, first = t & 1, t == 0
        f = np.empty((19, N))
        for q in range(19):
            up_p, up_i = up[q]
            if P == 0:
                f[q] = np.where(fluid, pop[q, np.where(fluid, self_i, 0)], 0.0)
            else:
                f[q] = np.where(up_p, pop[OPP[q], np.where(up_p, up_i, 0)], 0.0)
            if not first:
                f[q] = np.where(up_p, f[q], olbm.W32[q])
        val = A.copy()
        if not first:
            val[l_hasv] += w2[l_hasv] * B[P ^ 1][l_v[l_hasv]]
        f[l_t[l_ovr], l_node[l_ovr]] = val[l_ovr]
        fp, rho, u, v, w = olbm.macro_collide("bgk", f, (~fluid).astype(np.uint8), om1=om)
        assert np.isfinite(fp[:, fluid]).all()
        macros = (rho, u, v, w)
        for q in range(19):
            if P == 0:
                pop[OPP[q], self_i[fluid]] = fp[q, fluid]
                o = oob(q) & fluid
                pop[OPP[q], up[q][1][o]] = fp[q, o]
            else:
                dn_p, dn_i = up[OPP[q]]  # x + c_q
                wsel = fluid & dn_p
                pop[q, dn_i[wsel]] = fp[q, wsel]
                o = oob(q) & fluid
                pop[q, self_i[o]] = fp[q, o]
        A = w1 * fp[l_opp, l_node] + np.where(l_hasv, 0.0, w2 * fp[l_t, l_node])
        B[P][tb["ff_idx"]] = fp[f_p, f_node]
    return macros


def test_storage_records_count_present_nodes():
    """build_storage: present = fluid, ghost and plane-feeding solid nodes; the
    storage index of every present node is its rank in padded C order."""
    c = cases.config1()
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    bcs = [dict(axis=0, plane=0), dict(axis=0, plane=nx - 1)]
    sb = build_storage(nx, ny, nz, c["solid_flow"], (0, 0, 0), bcs)
    pres = sb["present"]
    assert pres.shape == (nx + 2, ny + 2, nz + 2)
    inner = pres[1:-1, 1:-1, 1:-1]
    assert np.all(inner[~c["solid_flow"]])
    # walls next to the inlet/outlet planes feed plane nodes: they keep slots
    assert inner[0][c["solid_flow"][0]].any() and not inner[nx // 2][c["solid_flow"][nx // 2]].any()
    I, J, K = (a.astype(np.int64) for a in np.nonzero(np.ones_like(pres)))
    p, idx = _locate(sb, I, J, K)
    assert np.array_equal(p, pres.reshape(-1))
    assert np.array_equal(idx[p], np.arange(sb["n_store"]))


@pytest.mark.parametrize("slot_lanes", [False, True])
@pytest.mark.parametrize("case", ["periodic_pack", "walled_box"])
def test_compact_storage_model_matches_oracle(case, slot_lanes):
    if case == "periodic_pack":
        c = cases.periodic_pack(n=24, seed=3)
        per, solid = (1, 1, 1), c["solid"]
    else:
        c = cases.config1()
        per, solid = (0, 0, 0), c["solid_flow"]
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    N = nx * ny * nz
    rng = np.random.default_rng(0)
    f0 = olbm.feq(1 + 0.01 * rng.random(N), *(0.02 * rng.standard_normal((3, N))), np.float64)
    om = 1.0 / 0.55
    steps = 7
    mm = model_run(nx, ny, nz, solid, c["links"], per, f0, steps, om, compact=False, storage="compact",
                   slot_lanes=slot_lanes)
    f = f0.copy()
    for _ in range(steps):
        f, *mo = olbm.step(f, solid, nx, ny, nz, per, "bgk", om, 1.0, None, 0.0, c["links"])
    for a, b in zip(mm, mo):
        assert np.abs(a - b).max() < 1e-12


@pytest.mark.parametrize("compact", [True, False])
@pytest.mark.parametrize("case", ["periodic_pack", "walled_box"])
def test_aa_model_matches_oracle(case, compact):
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
    mm = model_run(nx, ny, nz, solid, c["links"], per, f0, steps, om, compact=compact)
    f = f0.copy()
    for _ in range(steps):
        f, *mo = olbm.step(f, solid, nx, ny, nz, per, "bgk", om, 1.0, None, 0.0, c["links"])
    for a, b in zip(mm, mo):
        assert np.abs(a - b).max() < 1e-12


@pytest.mark.parametrize("compact", [True, False])
def test_entry_codes_match_node_masks(compact):
    """The fused step's per-entry codes (links, solid sources, publications, with
    per-wavefront first-entry offsets over the compacted fluid-node list) describe
    exactly the per-node masks used by the plane kernels (sp_mask)."""
    c = cases.periodic_pack(n=24, seed=3)
    n = c["nx"]
    tb = build_tables(n, n, n, c["solid"], c["links"], (1, 1, 1), compact=compact)
    wm = tb["wave_meta"].astype(np.int64)
    waves = (tb["n_lanes"] + 63) // 64
    assert wm.shape == (waves + 1, 4)
    fl = tb["lane_nodes"].astype(np.int64)
    assert np.array_equal(fl, np.nonzero(~c["solid"].reshape(-1))[0] if compact else np.arange(n ** 3))
    assert tuple(wm[-1]) == (tb["n_links"], tb["n_v"], tb["ff_node"].size, tb["n_src"])
    N = n ** 3
    L, S, V, F = (np.zeros(N, np.int64) for _ in range(4))
    lk = tb["lk_code"].astype(np.int64)
    node = fl[_group_of(wm[:, 0], lk.size) * 64 + (lk & 63)]
    assert np.array_equal(node, tb["link_node"])
    assert np.array_equal((lk >> 6) & 31, tb["link_dir"]) and np.array_equal((lk >> 11) & 31, OPP[tb["link_dir"]])
    np.bitwise_or.at(L, node, 1 << ((lk >> 6) & 31))
    hv = ((lk >> 16) & 1).astype(bool)
    np.bitwise_or.at(V, node[hv], 1 << ((lk[hv] >> 6) & 31))
    v_idx = wm[_group_of(wm[:, 0], lk.size), 1] + (lk >> 18)
    assert np.array_equal(v_idx[hv], np.arange(tb["n_v"]))
    sc = tb["s_code"].astype(np.int64)
    np.bitwise_or.at(S, fl[_group_of(wm[:, 3], sc.size) * 64 + (sc & 63)], 1 << ((sc >> 6) & 31))
    if not compact:  # solid lanes write w_q themselves: no substitution entries
        assert sc.size == 0 and tb["n_src"] == 0
        S[tb["sp_idx"]] = tb["sp_mask"].view(np.uint32).astype(np.int64)[:, 1]
    fc = tb["f_code"].astype(np.int64)
    f_node = fl[_group_of(wm[:, 2], fc.size) * 64 + (fc & 63)]
    assert np.array_equal(f_node, tb["ff_node"])
    # the publication's B slot rides in the code's upper bits, relative to its group's first slot
    assert tb["ff_packed"]
    assert np.array_equal(wm[_group_of(wm[:, 2], fc.size), 1] + (fc >> 11), tb["ff_idx"])
    # w2 = 1 - w1 bit for bit (the kernel forms it instead of loading it)
    assert np.array_equal(tb["w2"], np.float32(1.0) - tb["w1"])
    np.bitwise_or.at(F, f_node, 1 << ((fc >> 6) & 31))
    sp = tb["sp_idx"]
    masks = tb["sp_mask"].view(np.uint32).astype(np.int64)
    for col, arr in enumerate((L, S, V, F)):
        assert np.array_equal(arr[sp], masks[:, col]), col
        assert not np.any(np.delete(arr, sp))


def test_slot_lane_tables():
    """Compact storage with one lane per node owning slots: the lanes are the fluid
    nodes and the solid nodes feeding the inlet/outlet planes, in node order; there
    is no source list, and the link and publication entries decode to the same nodes
    and directions as with one lane per node."""
    c = cases.config1()
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    bcs = [dict(axis=0, plane=0), dict(axis=0, plane=nx - 1)]
    sb = build_storage(nx, ny, nz, c["solid_flow"], (0, 0, 0), bcs)
    mask = sb["present"][1:-1, 1:-1, 1:-1]
    tb = build_tables(nx, ny, nz, c["solid_flow"], c["links"], (0, 0, 0), bcs=bcs, compact=True, lane_mask=mask)
    ref = build_tables(nx, ny, nz, c["solid_flow"], c["links"], (0, 0, 0), bcs=bcs, compact=False)
    fl = tb["lane_nodes"].astype(np.int64)
    assert np.array_equal(fl, np.nonzero(mask.reshape(-1))[0])
    assert (c["solid_flow"].reshape(-1)[fl]).any()  # plane-feeding solid nodes have lanes
    assert tb["n_src"] == 0 and tb["s_code"].size == 0
    for t in (tb, ref):
        wm = t["wave_meta"].astype(np.int64)
        lk = t["lk_code"].astype(np.int64)
        lanes = t["lane_nodes"].astype(np.int64)
        t["_nodes"] = lanes[_group_of(wm[:, 0], lk.size) * 64 + (lk & 63)]
        t["_dirs"] = (lk >> 6) & 31
    assert np.array_equal(tb["_nodes"], tb["link_node"])
    key = lambda t: np.sort(t["_nodes"] * 19 + t["_dirs"])  # noqa: E731
    assert np.array_equal(key(tb), key(ref))
    with pytest.raises(ValueError):
        build_tables(nx, ny, nz, c["solid_flow"], c["links"], (0, 0, 0), compact=True,
                     lane_mask=np.zeros_like(mask))


@pytest.mark.parametrize("compact", [True, False])
def test_link_order_and_node_addressing(compact):
    """Within every wavefront the links with a second node come first (entry
    e < nV reads B slot v_lo + e), and the per-node addressing of the load,
    materialize and plane kernels (plain links from sp_base.x, second-node
    links from sp_base.w, by popcount of the lower mask bits) finds each link."""
    c = cases.periodic_pack(n=24, seed=3)
    n = c["nx"]
    tb = build_tables(n, n, n, c["solid"], c["links"], (1, 1, 1), compact=compact)
    wm = tb["wave_meta"].astype(np.int64)
    lk = tb["lk_code"].astype(np.int64)
    grp = _group_of(wm[:, 0], lk.size)
    e = np.arange(lk.size) - wm[grp, 0]
    hv = ((lk >> 16) & 1).astype(bool)
    assert hv.any() and (~hv).any()
    assert np.array_equal(e < np.diff(wm[:, 1])[grp], hv)
    assert np.array_equal((lk >> 18)[hv], e[hv])
    masks = tb["sp_mask"].view(np.uint32).astype(np.int64)
    base = tb["sp_base"].astype(np.int64)
    found = 0
    for row, x in enumerate(tb["sp_idx"]):
        Lm, Vm = int(masks[row, 0]), int(masks[row, 2])
        for q in range(19):
            if not (Lm >> q) & 1:
                continue
            below = (1 << q) - 1
            if (Vm >> q) & 1:
                li = base[row, 3] + bin(Vm & below).count("1")
                assert base[row, 1] + bin(Vm & below).count("1") == np.cumsum(hv)[li] - 1
            else:
                li = base[row, 0] + bin(Lm & ~Vm & below).count("1")
            assert tb["link_node"][li] == x and tb["link_dir"][li] == q
            found += 1
    assert found == tb["n_links"]
