"""Analytic pins of the build extensions without a reference counterpart
(parity unpinned against the reference, pinned by physics):

* Guo body force + textbook Bouzidi + fully periodic AA stepping: plane
  Poiseuille flow between walls at arbitrary positions (q = 0.3 and 0.7),
  u(y) = F / (2 nu) (y - y1)(y2 - y).
* FP16 / BF16 population storage (deviations from w_i): flow fields against
  the FP64 oracle of the same reference step.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fluxcore3d_amd.solver.d3q19 import mrt_rates  # noqa: E402
from oracle import lbm as olbm  # noqa: E402
from tests import cases  # noqa: E402

DEV = "cuda"


def _plates(ny=24, y1=0.7, y2=None):
    nx, nz = 4, 4
    y2 = ny - 1 - 0.3 if y2 is None else y2
    j = np.arange(ny, dtype=np.float64)
    d = np.minimum(j - y1, y2 - j)
    phi = np.broadcast_to(d[None, :, None], (nx, ny, nz)).astype(np.float32).copy()
    return phi, phi < 0, (nx, ny, nz), y1, y2


@pytest.mark.parametrize("kind", ["bgk", "trt", "mrt"])
def test_poiseuille_textbook_bouzidi_guo_force(kind):
    from fluxcore3d_amd.solver.aa_engine import AAEngine
    phi, solid, (nx, ny, nz), y1, y2 = _plates()
    # the channel is periodic in x and z: diagonal wall links wrap across those faces
    # (phi does not vary along x or z, so the q search is unaffected by the clamp)
    links = olbm.ibb_from_sdf(solid, phi, periodic=(True, True, True))
    tau = 0.8
    om = 1.0 / tau
    nu = (tau - 0.5) / 3.0
    # u_max = 6e-3: large enough that FP32 rounding of the populations (f ~ 1/18) stays
    # well below the discretisation error measured here; Re ~ 1, a linear regime in which
    # the relative error does not depend on Fx
    Fx = 1e-5
    om_minus = 1.0 / (0.5 + 0.1875 / (tau - 0.5))
    eng = AAEngine(nx, ny, nz, solid, periodic=(1, 1, 1), collision=kind, omega=om, omega_minus=om_minus,
                   S=mrt_rates(om) if kind == "mrt" else None, force=(Fx, 0, 0), links=links, ibb_mode="textbook",
                   device=DEV)
    f0 = np.broadcast_to(olbm.W32[:, None], (19, nx * ny * nz)).copy()
    eng.load(torch.as_tensor(f0, device=DEV))
    N = nx * ny * nz
    macro = [torch.empty(N, device=DEV) for _ in range(4)]
    for _ in range(12000):
        eng.step(*macro)
    u = macro[1].cpu().numpy().reshape(nx, ny, nz)[1, :, 1]
    y = np.arange(ny, dtype=np.float64)
    fluid = ~solid[1, :, 1]
    exact = Fx / (2 * nu) * (y - y1) * (y2 - y)
    err = np.linalg.norm(u[fluid] - exact[fluid]) / np.linalg.norm(exact[fluid])
    # a float64 model of the same data flow gives 1.4e-3 (BGK) and 1.6e-3 (TRT) at this resolution
    assert err < 5e-3, err
    # Interpolated bounce-back is not mass-conserving for q != 1/2 (walls at q = 0.3 and
    # 0.7 here), so a small transverse velocity remains: observed 1.3e-3 of the axial
    # maximum for TRT, below 1e-3 for BGK and MRT.
    assert np.abs(macro[2].cpu().numpy()).max() < 2e-3 * exact.max()


def test_reference_and_textbook_bouzidi_differ_as_documented():
    """At q = 1/2 textbook BFL is halfway bounce-back (no-slip); the reference's
    direction choice (lbm_kernels.cu:567-587) is not, so its wall slips."""
    from fluxcore3d_amd.solver.aa_engine import AAEngine
    phi, solid, (nx, ny, nz), _, _ = _plates(y1=0.5, y2=22.5)
    links = olbm.ibb_from_sdf(solid, phi, periodic=(True, True, True))
    out = {}
    for mode in ("reference", "textbook"):
        eng = AAEngine(nx, ny, nz, solid, periodic=(1, 1, 1), collision="bgk", omega=1.0, force=(1e-6, 0, 0),
                       links=links, ibb_mode=mode, device=DEV)
        eng.load(torch.as_tensor(np.broadcast_to(olbm.W32[:, None], (19, nx * ny * nz)).copy(), device=DEV))
        macro = [torch.empty(nx * ny * nz, device=DEV) for _ in range(4)]
        for _ in range(3000):
            eng.step(*macro)
        out[mode] = macro[1].cpu().numpy().reshape(nx, ny, nz)[1, :, 1]
    near_wall = out["textbook"][1] / out["textbook"][ny // 2]
    assert near_wall < 0.15  # first fluid node of a no-slip channel
    assert np.abs(out["reference"] - out["textbook"]).max() > 0.05 * out["textbook"].max()


@pytest.mark.parametrize("geom", ["pack", "gyroid"])
@pytest.mark.parametrize("storage,tol", [("fp32", None), ("fp16", 1e-3), ("bf16", 1e-2)])
def test_storage_precision_vs_fp64_oracle(storage, tol, geom):
    """FP32 / FP16 / BF16 population storage against the FP64 oracle on a sphere
    pack (40^3) and on one period of the gyroid precision workload (48^3), 100
    MRT steps from a random velocity field (errors relative to the largest speed).

    FP32: every one of the 100 GPU steps is compared with one FP32 oracle step from
    the GPU's own state (1e-5; the MRT collision is evaluated in moment space on the
    GPU and through the dense inverse in the oracle, so the two round differently).
    Over the 100 steps the random field decays and FP32 rounding differences grow
    (the FP32 oracle itself ends 4e-5 from FP64 on the gyroid), so the accumulated
    FP32 error is judged against the FP32 oracle's own distance from FP64: the GPU
    must be no further from FP64 than the oracle's FP32 arithmetic is."""
    from fluxcore3d_amd.solver.aa_engine import AAEngine
    c = cases.periodic_pack(n=40, seed=5) if geom == "pack" else cases.gyroid_box()
    n = c["nx"]
    om = 1.0 / 0.55
    S = mrt_rates(om, s_e=1.2, s_bulk=1.2, s_ghost=1.2)
    rng = np.random.default_rng(0)
    N = n ** 3
    rho = 1 + 0.005 * rng.random(N)
    u = 0.02 * rng.standard_normal((3, N))
    f0 = olbm.feq(rho, *u, np.float64)
    eng = AAEngine(n, n, n, c["solid"], periodic=(1, 1, 1), collision="mrt", omega=om, S=S, links=c["links"],
                   storage=storage, device=DEV)
    eng.load(torch.as_tensor(f0.astype(np.float32), device=DEV))
    macro = [torch.empty(N, device=DEV) for _ in range(4)]
    steps = 100
    f = f0.copy()
    step_err = 0.0
    for _ in range(steps):
        if storage == "fp32":
            f_gpu = eng.materialize().cpu().numpy().reshape(19, -1)
            _, *m1 = olbm.step(f_gpu, c["solid"], n, n, n, (1, 1, 1), "mrt", om, 1.0, S, 0.0, c["links"])
        eng.step(*macro)
        f, rho_w, u_w, v_w, w_w = olbm.step(f, c["solid"], n, n, n, (1, 1, 1), "mrt", om, 1.0, S.astype(np.float64),
                                            0.0, c["links"])
        if storage == "fp32":
            um = np.abs(np.stack(m1[1:])).max()
            step_err = max(step_err, max(np.abs(macro[a + 1].cpu().numpy() - m1[a + 1]).max() for a in range(3)) / um)
    umax = np.abs(np.stack([u_w, v_w, w_w])).max()
    err = max(np.abs(macro[a + 1].cpu().numpy() - (u_w, v_w, w_w)[a]).max() for a in range(3)) / umax
    if storage != "fp32":
        assert err < tol, (storage, err)
        return
    assert step_err <= 1e-5, step_err
    f32 = f0.astype(np.float32)
    for _ in range(steps):
        f32, *m32 = olbm.step(f32, c["solid"], n, n, n, (1, 1, 1), "mrt", om, 1.0, S, 0.0, c["links"])
    err_o32 = max(np.abs(m32[a + 1] - (u_w, v_w, w_w)[a]).max() for a in range(3)) / umax
    assert err <= max(1.5 * err_o32, 1e-5), (err, err_o32)
