"""GPU parity of the multigrid kernels and solvers against oracle/mg.py."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import mg  # noqa: E402
from tests import thermal_cases  # noqa: E402

DEV = "cuda"


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _fields(shape, seed=0):
    rng = np.random.default_rng(seed)
    k = (0.1 + rng.random(shape)).astype(np.float32)
    rcp = (0.5 + rng.random(shape)).astype(np.float32)
    u, v, w = ((0.05 * rng.standard_normal(shape)).astype(np.float32) for _ in range(3))
    return k, rcp, u, v, w


@pytest.mark.parametrize("ax,sign", [(0, 1), (1, -1), (2, 1)])
def test_coefficients_bit_exact_and_mop(ax, sign):
    from fluxcore3d_amd.solver.mg_cuda_ext import MG_CUDA
    from fluxcore3d_amd.solver.nlevel_mg_cuda import NLevelGeometricMGSolver
    shape = (12, 10, 14)
    k, rcp, u, v, w = _fields(shape, ax)
    mask = np.ones(shape, bool)
    s = NLevelGeometricMGSolver(torch.device(DEV), torch.float32, ax, sign, 1.5, omega=0.6)
    s.build_if_needed(T(k), T(rcp), T(u), T(v), T(w), T(mask), 0.3)
    out_idx = shape[ax] - 1 if sign == 1 else 0
    co = mg.build_coeffs_full(k, rcp, u, v, w, ax, out_idx, 1.0, 1.5)
    lv = s._levels[0]
    for name, ref in zip(("aE", "aW", "aN", "aS", "aU", "aD", "diagL"), co):
        np.testing.assert_array_equal(lv[name].cpu().numpy(), ref, err_msg=name)
    x = np.random.default_rng(1).standard_normal(shape).astype(np.float32)
    Lo = torch.empty(shape, device=DEV)
    MG_CUDA.mop_fused(T(x), *[lv[n] for n in ("aE", "aW", "aN", "aS", "aU", "aD", "diagL")], Lo, 0.3, ax, out_idx)
    assert rel(Lo.cpu().numpy(), mg.mop(x, co, 0.3, ax, out_idx)) < 1e-6
    xs = T(x)
    b = np.random.default_rng(2).standard_normal(shape).astype(np.float32)
    inv = lv["inv_diagM"]
    MG_CUDA.jacobi_smooth(xs, T(b), Lo, inv)
    want = x + inv.cpu().numpy() * (b - Lo.cpu().numpy())
    assert rel(xs.cpu().numpy(), want) < 1e-6


def _solver_pair(c, omega=0.6):
    from fluxcore3d_amd.solver.nlevel_mg_cuda import NLevelGeometricMGSolver
    s = NLevelGeometricMGSolver(torch.device(DEV), torch.float32, c["ax"], c["sign"], c["s"], omega=omega)
    s.build_if_needed(T(c["k"]), T(c["rcp"]), T(c["u"]), T(c["v"]), T(c["w"]), T(c["mask"]), c["dt"])
    h = mg.Hierarchy(c["k"], c["rcp"], c["u"], c["v"], c["w"], c["mask"], c["ax"], c["sign"], c["s"], omega=omega)
    h.set_dt(c["dt"])
    return s, h


def test_vcycle_matches_oracle():
    c = thermal_cases.sphere_conduction(64, 32, 32, dt=10.0)
    s, h = _solver_pair(c)
    assert [lv["shape"] for lv in s._levels] == [lv["shape"] for lv in h.levels]
    for a, b in zip(s._levels, h.levels):
        np.testing.assert_array_equal(a["mask_f"].cpu().numpy(), b["mask"])
        assert rel(a["inv_diagM"].cpu().numpy(), b["invD"]) < 1e-6
    r = c["b"].astype(np.float32)
    z = s.precondition(T(r).view(-1)).cpu().numpy()
    assert rel(z, h.precond(r)) < 1e-5


@pytest.mark.parametrize("solver", ["fcg", "gcr"])
def test_queued_iteration_leaves_the_solve_unchanged(solver):
    """FCG and GCR queue iteration k+1 before they read iteration k's stopping test
    (which runs on the device). Iterate, iteration count and history equal, bit for
    bit, those of the loop that waits for the test after every step (lookahead=False;
    an on_iter callback also selects it): at convergence, at max_it, and at a tolerance near the
    FP64 floor where the recursive residual can pass tol before the true one. The
    workspace is reused between calls, so a finished solve's flag must not leak."""
    from fluxcore3d_amd.solver.mg_pcg import FCGWorkspace, GCRWorkspace, fcg_fp64, gcr_fp64
    c = thermal_cases.sphere_conduction(64, 32, 32, dt=10.0)
    s, _ = _solver_pair(c)
    shape = s._levels[0]["shape"]
    if solver == "fcg":
        fn, kw, ws = fcg_fp64, {}, FCGWorkspace(shape, torch.device(DEV))
    else:
        fn, kw, ws = gcr_fp64, {"m": 6}, GCRWorkspace(shape, torch.device(DEV), 6)
    b = T(c["b"])
    for tol, max_it in ((1e-8, 200), (1e-8, 5), (1e-14, 60)):
        _, ahead = fn(s, b, tol=tol, max_it=max_it, ws=ws, lookahead=True, **kw)
        x_ahead = ahead["x64"].clone()
        _, wait = fn(s, b, tol=tol, max_it=max_it, ws=ws, lookahead=False, **kw)
        assert ahead["it"] == wait["it"] and ahead["converged"] == wait["converged"], (tol, max_it)
        assert ahead["hist"] == wait["hist"], (tol, max_it)
        assert torch.equal(x_ahead, wait["x64"]), (tol, max_it)
        if max_it == 200:
            assert ahead["converged"] and ahead["res_rel"] < tol
        if max_it == 5:
            assert ahead["it"] == 5 and not ahead["converged"]


def test_solve_cycles_match_oracle():
    c = thermal_cases.sphere_conduction(64, 32, 32, dt=10.0)
    s, h = _solver_pair(c)
    x, info = s.solve(T(c["b"]), max_cycles=6, tol=1e-30)
    xo, hist = h.solve(c["b"].astype(np.float32), max_cycles=6, tol=1e-30)
    assert info["it"] == len(hist) == 6
    assert rel(x.cpu().numpy(), xo) < 1e-5  # 4e-7 measured on an MI355X
    assert abs(info["res_rel"] - hist[-1]) / hist[-1] < 1e-5


def test_fcg_fp64_true_residual_and_solution():
    from fluxcore3d_amd.solver.mg_pcg import fcg_fp64
    c = thermal_cases.sphere_conduction(64, 32, 32, dt=1.0e6)
    s, h = _solver_pair(c)
    x, info = fcg_fp64(s, T(c["b"]), tol=1e-8, max_it=200)
    assert info["converged"], info["hist"][-5:]
    # independent FP64 residual of the returned solution
    co64 = tuple(a.astype(np.float64) for a in h.levels[0]["co"])
    x64 = info["x64"].cpu().numpy()
    m = c["mask"].astype(np.float64)
    r = (c["b"].astype(np.float64) * m - mg.mop(x64, co64, c["dt"], 0, h.levels[0]["out_idx"])) * m
    assert np.linalg.norm(r) / np.linalg.norm(c["b"] * m) < 1.01e-8
    xo, oinfo = mg.fcg_mg_fp64(h, c["b"], tol=1e-8, max_it=200)
    assert rel(x64 * m, xo) < 1e-6
    assert abs(info["it"] - oinfo["it"]) <= 2


@pytest.mark.parametrize("m", [2, 4])
def test_gcr_fp64_monotone_and_solution(m):
    """Flexible GCR(m) with the V-cycle: the residual never grows, the FP64 true
    residual of the returned x is below the tolerance (checked independently with
    the oracle operator), and x equals the oracle FCG solution."""
    from fluxcore3d_amd.solver.mg_pcg import gcr_fp64
    c = thermal_cases.sphere_conduction(64, 32, 32, dt=1.0e6)
    s, h = _solver_pair(c)
    x, info = gcr_fp64(s, T(c["b"]), tol=1e-8, max_it=200, m=m)
    assert info["converged"], info["hist"][-5:]
    hist = np.asarray(info["hist"][:-1])  # the last entry is the true residual
    assert np.all(np.diff(hist) <= 1e-12 * hist[:-1] + 1e-300)
    co64 = tuple(a.astype(np.float64) for a in h.levels[0]["co"])
    x64 = info["x64"].cpu().numpy()
    msk = c["mask"].astype(np.float64)
    r = (c["b"].astype(np.float64) * msk - mg.mop(x64, co64, c["dt"], 0, h.levels[0]["out_idx"])) * msk
    assert np.linalg.norm(r) / np.linalg.norm(c["b"] * msk) < 1.01e-8
    xo, _ = mg.fcg_mg_fp64(h, c["b"], tol=1e-10, max_it=300)
    assert rel(x64 * msk, xo) < 1e-6


@pytest.mark.parametrize("ax,sign", [(0, 1), (1, -1), (2, 1)])
def test_face_form_same_bits_as_per_cell_arrays(ax, sign):
    """Conduction with uniform rho*c_p and no advection: every level takes the face
    form (3 arrays instead of 7), and the V-cycle, the V-cycle iteration, flexible
    CG and flexible GCR give the same bits and iteration counts as with the
    per-cell arrays. sign=+1 puts the outlet at the high end of the flow axis, so
    the low face there is the self ghost, held in the face array's last plane."""
    from fluxcore3d_amd.solver.mg_pcg import fcg_fp64, gcr_fp64
    from fluxcore3d_amd.solver.nlevel_mg_cuda import NLevelGeometricMGSolver
    c = dict(thermal_cases.sphere_conduction(64, 32, 32, dt=1.0e6), ax=ax, sign=sign)

    def make(faces):
        s = NLevelGeometricMGSolver(torch.device(DEV), torch.float32, ax, sign, c["s"], omega=0.6,
                                    face_coefficients=faces)
        s.build_if_needed(T(c["k"]), T(c["rcp"]), T(c["u"]), T(c["v"]), T(c["w"]), T(c["mask"]), c["dt"])
        return s

    sf, sc = make(True), make(False)
    assert all(lv["faces"] is not None for lv in sf._levels)
    assert all(lv["faces"] is None for lv in sc._levels)
    r = T(c["b"].astype(np.float32)).view(-1)
    assert torch.equal(sf.precondition(r).clone(), sc.precondition(r).clone())
    xf, fi = sf.solve(T(c["b"]), max_cycles=4, tol=1e-30)
    xc, ci = sc.solve(T(c["b"]), max_cycles=4, tol=1e-30)
    assert torch.equal(xf, xc) and fi["res_rel"] == ci["res_rel"]
    for solver in (fcg_fp64, gcr_fp64):
        _, fi = solver(sf, T(c["b"]), tol=1e-8, max_it=200)
        _, ci = solver(sc, T(c["b"]), tol=1e-8, max_it=200)
        assert fi["converged"] and fi["it"] == ci["it"] and fi["hist"] == ci["hist"], solver.__name__
        assert torch.equal(fi["x64"], ci["x64"]), solver.__name__


def test_face_form_declined_with_advection_or_varying_heat_capacity():
    """Advection or a varying rho*c_p makes the per-cell coefficients of neighbouring
    cells differ: the build keeps those levels on the seven per-cell arrays, and the
    V-cycle still matches the oracle."""
    base = thermal_cases.sphere_conduction(64, 32, 32, dt=10.0)
    rng = np.random.default_rng(5)
    for change in (dict(u=(0.05 * rng.standard_normal(base["k"].shape)).astype(np.float32)),
                   dict(rcp=(0.5 + rng.random(base["k"].shape)).astype(np.float32))):
        c = dict(base, **change)
        s, h = _solver_pair(c)
        assert s._levels[0]["faces"] is None
        r = c["b"].astype(np.float32)
        assert rel(s.precondition(T(r).view(-1)).cpu().numpy(), h.precond(r)) < 1e-5


@pytest.mark.parametrize("shape", [(64, 32, 32), (44, 36, 30)])
def test_fused_first_sweeps_and_restriction_bits(shape):
    """The V-cycle's first two sweeps from x = 0 run as one pass
    (fc_mg_sweep2_from_zero): the V-cycle gives the same bits as with two
    separate sweeps, also with odd extents (44x36x30 -> 22x18x15 -> 11x9x7) and
    through the caller's output buffer. The residual + restriction kernel equals, bit for bit,
    the reference's sequence: Lo from mop_fused, r = (b - Lo) * mask, the 2x2x2 sum
    in the order (di, dj, dk) = (0,0,0), (0,0,1), ... (nlevel_mg_cuda.py:289-295,
    521-528), / 8."""
    from fluxcore3d_amd import _lib
    from fluxcore3d_amd.solver.nlevel_mg_cuda import NLevelGeometricMGSolver
    c = thermal_cases.sphere_conduction(*shape, dt=10.0)

    def make(fused, faces=True):
        s = NLevelGeometricMGSolver(torch.device(DEV), torch.float32, c["ax"], c["sign"], c["s"], omega=0.6,
                                    face_coefficients=faces, fused_first_sweeps=fused)
        s.build_if_needed(T(c["k"]), T(c["rcp"]), T(c["u"]), T(c["v"]), T(c["w"]), T(c["mask"]), c["dt"])
        return s

    r = T(np.random.default_rng(3).standard_normal(c["k"].shape).astype(np.float32) * c["mask"]).view(-1)
    sa, sb = make(True), make(False)
    za = sa.precondition(r, masked_input=True).clone()
    zb = sb.precondition(r, masked_input=True).clone()
    assert torch.equal(za, zb)
    out = torch.full_like(r, float("nan"))
    assert torch.equal(sa.precondition(r, out=out, masked_input=True).reshape(-1), zb.reshape(-1))
    assert torch.equal(out, zb.reshape(-1))
    xa, ia = sa.solve(T(c["b"]), max_cycles=3, tol=1e-30)
    xb, ib = sb.solve(T(c["b"]), max_cycles=3, tol=1e-30)
    assert torch.equal(xa, xb) and ia["res_rel"] == ib["res_rel"]
    for faces in (True, False):
        s = make(True, faces)
        for level in range(s._nlevels - 1):
            lv, lc = s._levels[level], s._levels[level + 1]
            nx, ny, nz = lv["shape"]
            rng = np.random.default_rng(level)
            x = T(rng.standard_normal(lv["shape"]).astype(np.float32))
            b = T(rng.standard_normal(lv["shape"]).astype(np.float32))
            bc = torch.empty(lc["shape"], device=DEV)
            _lib.call("fc_mg_residual_restrict", x.data_ptr(), b.data_ptr(), lv["mask_f"].data_ptr(), lv["coef"],
                      float(s._dt_cache), nx, ny, nz, s.ax, lv["out_idx"], bc.data_ptr(),
                      torch.cuda.current_stream().cuda_stream)
            Lo = torch.empty(lv["shape"], device=DEV)
            s._Mop_cuda(x, level, Lo)
            res = ((b - Lo) * lv["mask_f"]).cpu().numpy()
            cx, cy, cz = lc["shape"]
            want = np.zeros((cx, cy, cz), np.float32)
            for di in (0, 1):
                for dj in (0, 1):
                    for dk in (0, 1):
                        want += res[di:2 * cx:2, dj:2 * cy:2, dk:2 * cz:2]
            want /= np.float32(8.0)
            np.testing.assert_array_equal(bc.cpu().numpy(), want, err_msg=f"level {level} faces {faces}")


def test_prolongation_two_cells_per_thread_same_bits():
    """fc_mg_prolong_correct takes two fine cells per thread (8-byte accesses) when nz
    is even and x, mask are 8-byte aligned, else one. Both give the same bits: the
    same buffers at a 4-byte offset force the one-cell kernel."""
    from fluxcore3d_amd import _lib
    for shape in ((64, 32, 32), (44, 36, 30)):
        nx, ny, nz = shape
        cx, cy, cz = nx // 2, ny // 2, nz // 2
        rng = np.random.default_rng(7)
        xc = T(rng.standard_normal((cx, cy, cz)).astype(np.float32))
        x0 = rng.standard_normal(shape).astype(np.float32)
        m0 = (rng.random(shape) > 0.2).astype(np.float32)
        outs = []
        for off in (0, 1):
            xb = torch.zeros(nx * ny * nz + 2, device=DEV)
            mb = torch.zeros(nx * ny * nz + 2, device=DEV)
            xv, mv = xb[off:off + x0.size], mb[off:off + x0.size]
            xv.copy_(T(x0).view(-1))
            mv.copy_(T(m0).view(-1))
            _lib.call("fc_mg_prolong_correct", xc.data_ptr(), xv.data_ptr(), mv.data_ptr(), nx, ny, nz, cx, cy, cz,
                      0.75, None, torch.cuda.current_stream().cuda_stream)
            outs.append(xv.clone())
        assert torch.equal(outs[0], outs[1]), shape
        want = torch.nn.functional.interpolate(xc[None, None], size=shape, mode="trilinear", align_corners=False)[0, 0]
        ref = (T(x0) + 0.75 * want) * T(m0)
        assert rel(outs[0].view(shape).cpu().numpy(), ref.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("zero_guess", [0, 1])
def test_coarse_solve_lds_same_bits_as_sweeps(zero_guess):
    """The coarsest level's sweeps in one workgroup with the iterate in LDS equal,
    bit for bit, the same number of separate fc_mg_sweep launches."""
    from fluxcore3d_amd import _lib
    from fluxcore3d_amd.solver.nlevel_mg_cuda import NLevelGeometricMGSolver
    c = thermal_cases.sphere_conduction(64, 32, 32, dt=10.0)
    s = NLevelGeometricMGSolver(torch.device(DEV), torch.float32, c["ax"], c["sign"], c["s"], omega=0.6)
    s.build_if_needed(T(c["k"]), T(c["rcp"]), T(c["u"]), T(c["v"]), T(c["w"]), T(c["mask"]), c["dt"])
    lv = s._levels[-1]
    nx, ny, nz = lv["shape"]
    assert nx * ny * nz <= 4096
    rng = np.random.default_rng(11)
    x0 = T(rng.standard_normal(lv["shape"]).astype(np.float32))
    b = T(rng.standard_normal(lv["shape"]).astype(np.float32))
    st = torch.cuda.current_stream().cuda_stream
    dt = float(s._dt_cache)
    x, tmp = x0.clone(), torch.empty_like(x0)
    _lib.call("fc_mg_coarse_solve", x.data_ptr(), tmp.data_ptr(), b.data_ptr(), lv["coef"], dt, nx, ny, nz, s.ax,
              lv["out_idx"], 20, zero_guess, st)
    src, dst = x0.clone(), torch.empty_like(x0)
    for sweep in range(20):
        _lib.call("fc_mg_sweep", src.data_ptr(), dst.data_ptr(), b.data_ptr(), lv["coef"], dt, nx, ny, nz, s.ax,
                  lv["out_idx"], int(zero_guess and sweep == 0), st)
        src, dst = dst, src
    assert torch.equal(x, src)


def test_fgmres_arnoldi_step_against_fp64_operator():
    """One Arnoldi step (fc_fgmres_arnoldi) against an FP64 restatement: w = A z with
    the oracle operator on the FP32 coefficients, dot products with the basis, classical
    Gram-Schmidt, the FP32 copy and the squared norm."""
    import ctypes
    from fluxcore3d_amd import _lib
    from fluxcore3d_amd.solver.mg_pcg import FGMRESWorkspace, GCRStencil
    c = thermal_cases.sphere_conduction(64, 32, 32, dt=1.0e6)
    s, h = _solver_pair(c)
    lv0 = s._levels[0]
    shape = lv0["shape"]
    n = int(np.prod(shape))
    ws = FGMRESWorkspace(shape, torch.device(DEV), 4)
    rng = np.random.default_rng(21)
    m = c["mask"].astype(np.float64).reshape(-1)
    V0 = rng.standard_normal(n) * m
    V1 = rng.standard_normal(n) * m
    z = (rng.standard_normal(n) * m).astype(np.float32)
    ws.fcg.r.copy_(T(V0))
    ws.V[0].copy_(T(V1))
    ws.scal[ws.norm_slot[0]] = float(V0 @ V0)
    ws.scal[ws.norm_slot[1]] = float(V1 @ V1)
    faces = lv0["faces"]
    st = GCRStencil(*[lv0[k].data_ptr() for k in ("aE", "aW", "aN", "aS", "aU", "aD", "diagL")],
                    lv0["mask_f"].data_ptr(), float(s._dt_cache), *shape, int(s.ax), int(lv0["out_idx"]),
                    *[f.data_ptr() for f in faces])
    Vp = (ctypes.c_void_p * 5)(ws.fcg.r.data_ptr(), *[v.data_ptr() for v in ws.V])
    zt = T(z)
    _lib.call("fc_fgmres_arnoldi", zt.data_ptr(), ctypes.byref(st), Vp, 1, ws.v32.data_ptr(), ws.partial.data_ptr(),
              ws.scal.data_ptr(), torch.cuda.current_stream().cuda_stream)
    co64 = tuple(a.astype(np.float64) for a in h.levels[0]["co"])
    w = (mg.mop(z.astype(np.float64).reshape(shape), co64, c["dt"], 0, h.levels[0]["out_idx"]).reshape(-1)) * m
    d0, d1 = w @ V0, w @ V1
    v2 = w - d0 / (V0 @ V0) * V0 - d1 / (V1 @ V1) * V1
    sc = ws.scal.cpu().numpy()
    assert abs(sc[ws.dot_slot[0]] - d0) <= 1e-12 * np.linalg.norm(w) * np.linalg.norm(V0)
    assert abs(sc[ws.dot_slot[1]] - d1) <= 1e-12 * np.linalg.norm(w) * np.linalg.norm(V1)
    got = ws.V[1].cpu().numpy()
    assert np.abs(got - v2).max() <= 1e-12 * np.abs(w).max()
    assert abs(sc[ws.norm_slot[2]] - v2 @ v2) <= 1e-12 * (v2 @ v2)
    np.testing.assert_array_equal(ws.v32.cpu().numpy(), got.astype(np.float32))


@pytest.mark.parametrize("m", [4, 16])
def test_fgmres_fp64_converges_to_the_solution(m):
    """Restarted flexible GMRES(m) with the V-cycle: the residual estimates never grow
    within a cycle, the stored basis is orthogonal, the FP64 true residual of the
    returned x is below the tolerance (checked with the oracle operator) and x equals
    the oracle FCG solution."""
    from fluxcore3d_amd.solver.mg_pcg import FGMRESWorkspace, fgmres_fp64
    c = thermal_cases.sphere_conduction(64, 32, 32, dt=1.0e6)
    s, h = _solver_pair(c)
    ws = FGMRESWorkspace(s._levels[0]["shape"], torch.device(DEV), m)
    x, info = fgmres_fp64(s, T(c["b"]), tol=1e-8, max_it=300, m=m, ws=ws)
    assert info["converged"], info["hist"][-5:]
    hist = info["hist"]  # estimates, with the true residual after each cycle
    assert all(b <= a * 1.01 for a, b in zip(hist, hist[1:])), hist
    k = info["it"] - m * ((info["it"] - 1) // m)  # steps of the last cycle: V_1..V_k are stored
    if k >= 2:
        Vs = torch.stack([v / v.norm() for v in ws.V[:k]])
        G = (Vs @ Vs.T).cpu().numpy()
        assert np.abs(G - np.eye(k)).max() < 1e-8
    co64 = tuple(a.astype(np.float64) for a in h.levels[0]["co"])
    x64 = info["x64"].cpu().numpy()
    msk = c["mask"].astype(np.float64)
    r = (c["b"].astype(np.float64) * msk - mg.mop(x64, co64, c["dt"], 0, h.levels[0]["out_idx"])) * msk
    assert np.linalg.norm(r) / np.linalg.norm(c["b"] * msk) < 1.01e-8
    xo, _ = mg.fcg_mg_fp64(h, c["b"], tol=1e-10, max_it=300)
    assert rel(x64 * msk, xo) < 1e-6


def test_cht_advection_system_gcr_and_fgmres():
    """The coupled heat-sink conduction-advection system (pin-fin plate at 1/4 scale,
    48x48x96: upwind advection along z, rho*c_p varying 2000x, so non-symmetric with
    per-cell coefficients): GMG-GCR(6) and GMG-FGMRES(16) reach an FP64 residual of
    1e-8, checked with the oracle's coefficients and operator in FP64, and agree on
    the solution. Flexible CG is not asked to converge on this operator (at full size
    its residual grows without bound)."""
    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.solver.mg_pcg import fgmres_fp64, gcr_fp64
    from fluxcore3d_amd.solver.nlevel_mg_cuda import NLevelGeometricMGSolver
    c = geometry.cht_problem(scale=0.25, device=DEV)
    s = NLevelGeometricMGSolver(torch.device(DEV), torch.float32, c["ax"], c["sign"], c["s"], omega=0.8)
    s.build_if_needed(c["k"], c["rcp"], c["u"], c["v"], c["w"], c["mask"], c["dt"])
    assert s._levels[0]["faces"] is None
    _, gi = gcr_fp64(s, c["b"], tol=1e-8, max_it=300, m=6)
    assert gi["converged"]
    xg = gi["x64"].clone()
    del gi
    _, fi = fgmres_fp64(s, c["b"], tol=1e-8, max_it=300, m=16)
    assert fi["converged"]
    xf = fi["x64"].clone()
    assert float((xg - xf).norm() / xf.norm()) < 1e-6
    f32 = lambda t: t.cpu().numpy().astype(np.float32)  # noqa: E731
    out_idx = c["shape"][c["ax"]] - 1
    co = mg.build_coeffs_full(f32(c["k"]), f32(c["rcp"]), f32(c["u"]), f32(c["v"]), f32(c["w"]), c["ax"], out_idx,
                              1.0, c["s"])
    co64 = tuple(a.astype(np.float64) for a in co)
    msk = c["mask"].cpu().numpy().astype(np.float64)
    b = c["b"].cpu().numpy().astype(np.float64) * msk
    for x in (xg, xf):
        x64 = x.cpu().numpy()
        r = (b - mg.mop(x64, co64, c["dt"], c["ax"], out_idx)) * msk
        assert np.linalg.norm(r) / np.linalg.norm(b) < 1.01e-8


def test_bicgstab_against_oracle_and_patch_routing():
    """MG-preconditioned BiCGSTAB (mg_pcg.py:103-175) on the non-symmetric heat-sink
    system (1/4 scale): its first residuals equal the oracle's BiCGSTAB (FP64 vectors,
    the FP32 hierarchy) to 1e-4, it reaches an FP64 residual of 1e-8 (checked with the
    oracle's operator) with the GCR solution, and patch_mg_pcg routes solve() by the
    reference's heuristic pe = 0.1 max(diagL) (mg_pcg.py:250-275)."""
    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.solver.mg_pcg import bicgstab_mg, gcr_fp64, patch_mg_pcg, pcg_mg
    from fluxcore3d_amd.solver.nlevel_mg_cuda import NLevelGeometricMGSolver
    c = geometry.cht_problem(scale=0.25, device=DEV)
    s = NLevelGeometricMGSolver(torch.device(DEV), torch.float32, c["ax"], c["sign"], c["s"], omega=0.8)
    s.build_if_needed(c["k"], c["rcp"], c["u"], c["v"], c["w"], c["mask"], c["dt"])
    f32 = lambda t: t.cpu().numpy().astype(np.float32)  # noqa: E731
    h = mg.Hierarchy(f32(c["k"]), f32(c["rcp"]), f32(c["u"]), f32(c["v"]), f32(c["w"]), f32(c["mask"]) > 0,
                     c["ax"], c["sign"], c["s"], omega=0.8)
    h.set_dt(c["dt"])
    b = f32(c["b"])
    _, oi = mg.bicgstab_mg(h, b, tol=1e-30, max_it=4, outer=np.float64)
    _, gi4 = bicgstab_mg(s, c["b"], tol=1e-30, max_it=4)
    assert np.allclose(gi4["hist"], oi["hist"], rtol=1e-4, atol=0.0), (gi4["hist"], oi["hist"])
    _, bi = bicgstab_mg(s, c["b"], tol=1e-8, max_it=300)
    assert bi["converged"], bi["hist"][-3:]
    xb = bi["x64"].clone()
    _, gi = gcr_fp64(s, c["b"], tol=1e-8, max_it=300, m=6)
    assert float((xb - gi["x64"]).norm() / gi["x64"].norm()) < 1e-6
    co64 = tuple(a.astype(np.float64) for a in h.levels[0]["co"])
    msk = c["mask"].cpu().numpy().astype(np.float64)
    b64 = c["b"].cpu().numpy().astype(np.float64) * msk
    r = (b64 - mg.mop(xb.cpu().numpy(), co64, c["dt"], c["ax"], h.levels[0]["out_idx"])) * msk
    assert np.linalg.norm(r) / np.linalg.norm(b64) < 1.01e-8
    pe = float(s._levels[0]["diagL"].max()) * 0.1
    for thr, direct in ((pe * 0.5, bicgstab_mg), (pe * 2.0, pcg_mg)):
        patch_mg_pcg(s, pe_threshold=thr)
        xp, pinfo = s.solve(c["b"], max_cycles=6, tol=1e-30)
        xd, dinfo = direct(s, c["b"], tol=1e-30, max_it=6)
        assert torch.equal(xp, xd) and pinfo["res_rel"] == dinfo["res_rel"], direct.__name__


@pytest.mark.parametrize("solver", ["pcg_mg", "gcr", "fgmres"])
def test_initial_guess_on_dirichlet_rows_is_ignored(solver):
    """The reference's pcg_mg masks x0 before its first residual (mg_pcg.py:50-51):
    values of x0 on Dirichlet rows (mask 0) must not leak into the neighbouring
    unknowns through the stencil. The solve from a random x0 that is nonzero on the
    masked wall planes equals the solve from x0 * mask, and the oracle's FCG solution."""
    from fluxcore3d_amd.solver.mg_pcg import fcg_fp64, fgmres_fp64, gcr_fp64
    c = thermal_cases.sphere_conduction(32, 16, 16, dt=1.0e3)
    s, h = _solver_pair(c, omega=0.8)
    rng = np.random.default_rng(3)
    x0 = (10.0 * rng.standard_normal(c["k"].shape)).astype(np.float32)
    assert np.abs(x0[~c["mask"]]).min() > 0
    fn, kw = {"pcg_mg": (fcg_fp64, {}), "gcr": (gcr_fp64, {"m": 6}), "fgmres": (fgmres_fp64, {"m": 16})}[solver]
    _, a = fn(s, T(c["b"]), x0=T(x0), tol=1e-10, max_it=400, **kw)
    xa = a["x64"].clone()
    _, b = fn(s, T(c["b"]), x0=T(x0 * c["mask"]), tol=1e-10, max_it=400, **kw)
    assert a["converged"] and b["converged"]
    assert float((xa - b["x64"]).abs().max()) <= 1e-8 * float(b["x64"].abs().max())
    assert float(xa.view(-1)[torch.as_tensor(~c["mask"].reshape(-1), device=DEV)].abs().max()) == 0.0
    h.set_dt(c["dt"])
    xo, _ = mg.fcg_mg_fp64(h, c["b"].astype(np.float32), tol=1e-12, max_it=400)
    got = xa.cpu().numpy().reshape(c["k"].shape)
    assert np.abs(got - xo).max() <= 1e-7 * np.abs(xo).max()


def test_fgmres_exact_breakdown_with_identity_operator():
    """dt = 0 makes A = I. With the identity preconditioner and a right-hand side with
    one nonzero entry, every sum is exact: A z = r, h_00 = 1 and the second basis
    vector is exactly zero (an exact breakdown), while the step that uses it is
    already queued. The Givens kernel leaves that column empty, the host drops it,
    and the solve ends after one iteration with x = b exactly."""
    from fluxcore3d_amd.solver.mg_pcg import fgmres_fp64
    c = thermal_cases.sphere_conduction(16, 8, 8, dt=0.0)
    s, _ = _solver_pair(c)
    b = np.zeros(c["k"].shape, np.float32)
    b[5, 3, 4] = 1.5
    assert c["mask"][5, 3, 4]
    x, info = fgmres_fp64(s, T(b), tol=1e-12, max_it=16, m=16, precond=lambda v, z: z.copy_(v))
    x64 = info["x64"].cpu().numpy()
    assert np.isfinite(x64).all() and np.isfinite(info["hist"]).all()
    assert info["converged"] and info["it"] == 1, info
    np.testing.assert_array_equal(x64, b.astype(np.float64))


@pytest.mark.parametrize("n_steps", [1, 2, 3])
@pytest.mark.parametrize("level", [0, 1])
def test_chebyshev_smooth_matches_oracle(level, n_steps):
    """chebyshev_smooth (mg_pcg.py:195-247) on a level of the config-1 conduction
    hierarchy against the oracle's FP32 restatement (1e-6 of the largest value)."""
    from fluxcore3d_amd.solver.mg_pcg import chebyshev_smooth
    c = thermal_cases.sphere_conduction(64, 32, 32, dt=10.0)
    s, h = _solver_pair(c)
    shape = s._levels[level]["shape"]
    rng = np.random.default_rng(level * 10 + n_steps)
    x0 = rng.standard_normal(shape).astype(np.float32)
    b = rng.standard_normal(shape).astype(np.float32)
    lv = s._levels[level]
    lv["x"].copy_(T(x0).view_as(lv["x"]))
    lv["b_buf"].copy_(T(b).view_as(lv["b_buf"]))
    chebyshev_smooth(s, level, n_steps=n_steps)
    want = mg.chebyshev_smooth(h, level, x0, b, n_steps)
    got = lv["x"].cpu().numpy().reshape(shape)
    assert rel(got, want) < 1e-6, rel(got, want)
    assert np.abs(got - x0).max() > 1e-3  # it smoothed


@pytest.mark.parametrize("route", ["gcr", "fgmres"])
def test_patch_mg_pcg_nonsymmetric_route(route):
    """patch_mg_pcg(nonsymmetric_route=...): the heat-sink system with advection (1/4
    scale, per-cell coefficients) that the reference's heuristic sends to CG (here with
    a threshold above its pe) goes to the minimal-residual solver and reaches 1e-8; a
    symmetric conduction system (face form) keeps the reference's routing."""
    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.solver.mg_pcg import patch_mg_pcg
    from fluxcore3d_amd.solver.nlevel_mg_cuda import NLevelGeometricMGSolver
    c = geometry.cht_problem(scale=0.25, device=DEV)
    s = NLevelGeometricMGSolver(torch.device(DEV), torch.float32, c["ax"], c["sign"], c["s"], omega=0.8)
    s.build_if_needed(c["k"], c["rcp"], c["u"], c["v"], c["w"], c["mask"], c["dt"])
    assert s._levels[0].get("faces") is None
    pe = float(s._levels[0]["diagL"].max()) * 0.1
    patch_mg_pcg(s, pe_threshold=2.0 * pe)
    _, info = s.solve(c["b"], max_cycles=40, tol=1e-8)
    assert info["route"] == "pcg"
    patch_mg_pcg(s, pe_threshold=2.0 * pe, nonsymmetric_route=route)
    _, info = s.solve(c["b"], max_cycles=300, tol=1e-8)
    assert info["route"] == route and info["converged"] and info["res_rel"] < 1e-8, info["res_rel"]
    sym = thermal_cases.sphere_conduction(32, 16, 16, dt=1.0e3)
    s2, _ = _solver_pair(sym, omega=0.8)
    assert s2._levels[0].get("faces") is not None
    patch_mg_pcg(s2, pe_threshold=1e30, nonsymmetric_route=route)
    _, info2 = s2.solve(T(sym["b"]), max_cycles=100, tol=1e-8)
    assert info2["route"] == "pcg" and info2["converged"]
