"""The HIP path against the committed oracle fixtures (tests/golden/golden.npz,
written by tests/golden/make_golden.py; test_golden.py pins the oracle to them)."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tests.golden import make_golden  # noqa: E402

DEV = "cuda"
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")
LINK_KEYS = ("fluid_flat", "dir_i", "ff_flat", "ff_valid", "q", "q_ge_half", "w1", "w2")


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_links_and_q_match_fixture_bit_for_bit(golden):
    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.pre_post.sdf_tracing import bfl_q_and_weights, build_links_from_mask_gpu
    phi, solid, _ = geometry.sphere_channel()
    ff, di, _, fff, ok = build_links_from_mask_gpu(torch.as_tensor(solid, device=DEV))
    q, ge, w1, w2 = bfl_q_and_weights(torch.as_tensor(phi, device=DEV), (0, 0, 0), 1.0, ff, di)
    got = dict(fluid_flat=ff, dir_i=di, ff_flat=fff, ff_valid=ok, q=q, q_ge_half=ge, w1=w1, w2=w2)
    for k in LINK_KEYS:
        np.testing.assert_array_equal(got[k].cpu().numpy(), golden[f"c1_{k}"], err_msg=k)


def test_config1_steps_match_fixture(golden):
    """The fused step on the fixture's links: the z = 16 plane after 1, 10 and 100
    steps within 1e-5 of the largest speed (velocities; at least the inlet speed) and
    relative (density), and the total mass within 1e-6."""
    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.solver.aa_engine import AAEngine
    from fluxcore3d_amd.solver.d3q19 import W
    _, _, solid_flow = geometry.sphere_channel()
    nx, ny, nz = solid_flow.shape
    links = {k: golden[f"c1_{k}"] for k in LINK_KEYS if k != "q"}
    eng = AAEngine(nx, ny, nz, solid_flow, periodic=(0, 0, 0), collision="bgk", omega=1.0 / 0.6, links=links,
                   bcs=make_golden.config1_bcs(nx), device=DEV)
    n = nx * ny * nz
    eng.load(torch.as_tensor(np.broadcast_to(np.asarray(W, np.float32)[:, None], (19, n)).copy(), device=DEV))
    macro = [torch.empty(n, device=DEV) for _ in range(4)]
    for step in range(1, max(make_golden.STEPS) + 1):
        eng.step(*macro)
        if step not in make_golden.STEPS:
            continue
        planes = [m.view(nx, ny, nz)[:, :, make_golden.PLANE].cpu().numpy() for m in macro]
        want = [golden[f"c1_s{step}_{k}"] for k in ("rho", "u", "v", "w")]
        # the largest speed, at least the inlet's 0.05 (after one step the plane is still at rest)
        speed = max(float(np.sqrt(want[1] ** 2 + want[2] ** 2 + want[3] ** 2).max()), 0.05)
        assert np.abs(planes[0] - want[0]).max() / np.abs(want[0]).max() < 1e-5, step
        for g, w_ in zip(planes[1:], want[1:]):
            assert np.abs(g - w_).max() / speed < 1e-5, step
        mass = float(macro[0].double().sum())
        assert abs(mass - float(golden[f"c1_s{step}_mass"])) / mass < 1e-6, step


def test_multigrid_matches_fixture(golden):
    """One V-cycle against the FP32 and FP64 oracle V-cycles; GMG-FCG, GMG-GCR(6) and
    GMG-FGMRES(16) to an FP64 residual of 1e-10 against the oracle's FCG solution."""
    from fluxcore3d_amd.solver.mg_pcg import fcg_fp64, fgmres_fp64, gcr_fp64
    from fluxcore3d_amd.solver.nlevel_mg_cuda import NLevelGeometricMGSolver
    c = make_golden.mg_case()
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=DEV)  # noqa: E731
    s = NLevelGeometricMGSolver(torch.device(DEV), torch.float32, c["ax"], c["sign"], c["s"],
                                omega=make_golden.MG_OMEGA)
    s.build_if_needed(T(c["k"]), T(c["rcp"]), T(c["u"]), T(c["v"]), T(c["w"]), T(c["mask"]), c["dt"])
    assert len(s._levels) == 2
    z = s.precondition(T(golden["mg_r"]).view(-1), masked_input=True).cpu().numpy()
    scale = np.abs(golden["mg_z64"]).max()
    assert np.abs(z - golden["mg_z32"]).max() / scale < 1e-5
    assert np.abs(z - golden["mg_z64"]).max() / scale < 1e-4
    x64 = golden["mg_x64"]
    msk = c["mask"].astype(np.float64)
    for solver, kw in ((fcg_fp64, {}), (gcr_fp64, {"m": 6}), (fgmres_fp64, {"m": 16})):
        _, info = solver(s, T(c["b"]), tol=1e-10, max_it=400, **kw)
        assert info["converged"], solver.__name__
        got = info["x64"].cpu().numpy() * msk
        assert np.abs(got - x64).max() / np.abs(x64).max() < 1e-7, solver.__name__
