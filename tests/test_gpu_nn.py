"""Solver-side pieces of the learned preconditioner (no reference counterpart;
parity unpinned against the reference, pinned by the HIP operator and by
convergence of the FP64 true residual)."""
import numpy as np
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


def test_trained_network_one_step_reduction_on_held_out_problem():
    """What training buys, on a problem outside the training seeds: for errors of
    random spectra (the training distribution) the trained network's direction
    removes most of the scaled residual in one minimal-residual step
    (sin^2 between r~ and A~ z well below 1), the untrained network's almost none.
    (On the solver's own residuals the trained network does not beat the V-cycle;
    that is measured by bench.py's thermal leg.)"""
    import os
    from fluxcore3d_amd.solver import nn_precond as nn
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fluxcore3d_amd", "data",
                        "unet_precond.npz")
    trained = nn.UNetPrecond.load(path, device=DEV)
    untrained = nn.UNetPrecond(trained.channels, device=DEV, seed=11)
    samples, ops = nn.harvest_pairs(n=64, problems=1, seed0=7777, per_problem=6, device=DEV, log=None)

    def sin2(net):
        out = []
        with torch.no_grad():
            for x, e, m, pid, _ in samples:
                Az = ops[pid](net.forward_train(x[None].contiguous())[0, ..., 0] * m)
                r = ops[pid](e * m)
                c = float((Az * r).sum())
                out.append(1.0 - c * c / (float((Az * Az).sum()) * float((r * r).sum())))
        return float(np.mean(out))

    s_tr, s_un = sin2(trained), sin2(untrained)
    assert s_tr < 0.5 and s_un > 0.9, (s_tr, s_un)
