"""sdf_trilinear, bfl_q_from_sdf and bfl_weights_from_q (pre_post/sdf_tracing.py:88-205)
on the GPU against the oracle's FP32 restatement, bit for bit."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import lbm as olbm  # noqa: E402

DEV = "cuda"


def _phi(n=24, r=6.3, c=(11.4, 12.2, 10.9)):
    i, j, k = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    return (np.sqrt((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2) - r).astype(np.float32)


def test_sdf_trilinear_bit_exact():
    from fluxcore3d_amd.pre_post.sdf_tracing import sdf_trilinear
    rng = np.random.default_rng(4)
    phi = rng.standard_normal((9, 7, 11)).astype(np.float32)
    origin = np.array([0.3, -1.2, 2.0], np.float32)
    x = (origin + 0.5 * rng.uniform(-2.0, 12.0, (40, 25, 3))).astype(np.float32)  # in range and clamped
    got = sdf_trilinear(torch.as_tensor(phi, device=DEV), torch.as_tensor(origin), 0.5,
                        torch.as_tensor(x, device=DEV))
    assert tuple(got.shape) == (40, 25)
    np.testing.assert_array_equal(got.cpu().numpy(), olbm.sdf_points(phi, origin, 0.5, x))


def test_bfl_q_from_sdf_and_weights_bit_exact():
    from fluxcore3d_amd.pre_post.sdf_tracing import bfl_q_from_sdf, bfl_weights_from_q
    phi = _phi()
    rng = np.random.default_rng(6)
    n = 3000
    # half the rays start 0.2-2 outside the sphere and point roughly at its centre
    ctr = np.array([11.4, 12.2, 10.9])
    u = rng.standard_normal((n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    aimed = ctr + u * (6.3 + rng.uniform(0.2, 2.0, (n, 1)))
    x0 = np.where(np.arange(n)[:, None] % 2 == 0, aimed, rng.uniform(0.0, 23.0, (n, 3))).astype(np.float32)
    d = np.where(np.arange(n)[:, None] % 2 == 0, -u + 0.3 * rng.standard_normal((n, 3)), rng.standard_normal((n, 3)))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    L = rng.uniform(0.5, 3.0, n).astype(np.float32)
    T = lambda a: torch.as_tensor(a, device=DEV)  # noqa: E731
    for n_coarse, n_bisect in ((6, 10), (3, 4)):
        q = bfl_q_from_sdf(T(phi), torch.zeros(3), 1.0, T(x0), T(d), T(L), n_coarse, n_bisect)
        want = olbm.q_rays(phi, (0, 0, 0), 1.0, x0, d, L, n_coarse, n_bisect)
        np.testing.assert_array_equal(q.cpu().numpy(), want)
        assert 0.1 < float((want != 0.5).mean()) < 0.9  # both crossing and no-crossing rays
    ge, w1, w2 = bfl_weights_from_q(q)
    wge, ww1, ww2 = olbm.bfl_weights(want)
    np.testing.assert_array_equal(ge.cpu().numpy(), wge)
    np.testing.assert_array_equal(w1.cpu().numpy(), ww1)
    np.testing.assert_array_equal(w2.cpu().numpy(), ww2)
