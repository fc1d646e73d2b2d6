"""Solver-side pieces of the learned preconditioner (no reference counterpart;
parity unpinned against the reference, pinned by the HIP operator and by
convergence of the FP64 true residual)."""
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DEV = "cuda"


def _problem(n=32, seed=7, ratio=300.0):
    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.solver.nn_precond import _mg_for
    c = geometry.conduction_problem(n, seed=seed, ratio=ratio, device=DEV)
    return c, _mg_for(c, DEV)


def test_training_operator_matches_hip_operator():
    """The torch operator of the training loss, D^-1/2 A D^-1/2, equals the HIP
    FP64 operator (fc_mg_apply_A64) conjugated by the same diagonal scaling."""
    from fluxcore3d_amd.solver.nn_precond import ScaledOperator, apply_A64, jacobi_scale
    c, mg = _problem()
    op = ScaledOperator(mg)
    g = torch.Generator(device="cpu").manual_seed(0)
    u = torch.randn(op.shape, generator=g).to(DEV)
    dsq = jacobi_scale(mg).double()
    ref = (apply_A64(mg, u.reshape(-1).double() * dsq) * dsq).view(op.shape)
    got = op(u)
    assert (got.double() - ref).norm() <= 1e-5 * ref.norm()
    # the scaled diagonal is one on unmasked rows
    e0 = torch.zeros(op.shape, device=DEV)
    e0[5, 6, 7] = 1.0
    assert abs(float(op(e0)[5, 6, 7]) - 1.0) < 1e-5


def test_nn_fcg_reaches_tolerance_with_untrained_network():
    """An untrained network is a poor preconditioner; the driver must still
    reach the FP64 true-residual tolerance (through the GMG-PCG fallback)."""
    from fluxcore3d_amd.solver.mg_pcg import fcg_fp64
    from fluxcore3d_amd.solver.nn_precond import UNetPrecond, nn_fcg
    c, mg = _problem()
    net = UNetPrecond(channels=(8, 16, 16), device=DEV, seed=3)
    x, info = nn_fcg(mg, net, c["b"], c["k"], tol=1e-8, max_it=300)
    assert info["converged"] and info["res_rel"] < 1e-8
    xg, gi = fcg_fp64(mg, c["b"], tol=1e-10, max_it=300)
    assert (x - xg).norm() <= 1e-4 * xg.norm()
