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


def test_solve_cycles_match_oracle():
    c = thermal_cases.sphere_conduction(64, 32, 32, dt=10.0)
    s, h = _solver_pair(c)
    x, info = s.solve(T(c["b"]), max_cycles=6, tol=1e-30)
    xo, hist = h.solve(c["b"].astype(np.float32), max_cycles=6, tol=1e-30)
    assert info["it"] == len(hist) == 6
    assert rel(x.cpu().numpy(), xo) < 1e-4
    assert abs(info["res_rel"] - hist[-1]) / hist[-1] < 1e-3


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
