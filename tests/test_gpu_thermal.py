"""GPU checks of the thermal step kernels (behaviour of thermal_kernels.cu:31-189)
against closed forms computed with NumPy on the same inputs."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DEV = "cuda"


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV)


def test_rhs_and_theta0():
    from fluxcore3d_amd.solver.thermal_cuda_ext import THERMAL_CUDA
    rng = np.random.default_rng(0)
    shape = (9, 7, 5)
    To, T0, src, LT0 = (rng.standard_normal(shape).astype(np.float32) for _ in range(4))
    m = (rng.random(shape) < 0.8).astype(np.float32)
    rhs, th = torch.empty(shape, device=DEV), torch.empty(shape, device=DEV)
    THERMAL_CUDA.thermal_rhs(T(To), T(T0), T(src), T(LT0), T(m), rhs, th, 0.3)
    np.testing.assert_allclose(rhs.cpu().numpy(), m * ((To + np.float32(0.3) * src) - LT0), rtol=0, atol=1e-6)
    np.testing.assert_array_equal(th.cpu().numpy(), m * (To - T0))


@pytest.mark.parametrize("ax,out_idx", [(0, 8), (0, 0), (2, 4)])
def test_reconstruct_neumann_faces(ax, out_idx):
    from fluxcore3d_amd.solver.thermal_cuda_ext import THERMAL_CUDA
    rng = np.random.default_rng(1)
    nx, ny, nz = 9, 7, 5
    T0, th = (rng.standard_normal((nx, ny, nz)).astype(np.float32) for _ in range(2))
    got = THERMAL_CUDA.thermal_reconstruct(T(T0), T(th), nx, ny, nz, ax, out_idx).cpu().numpy()
    val = T0 + th
    want = val.copy()
    n = (nx, ny, nz)
    for c in np.ndindex(nx, ny, nz):
        src = list(c)
        for a in range(3):  # face priority x, y, z
            if c[a] == n[a] - 1:
                src[a] = c[a] - 1
                break
            if c[a] == 0 and (a != ax or out_idx == 0):
                src[a] = 1
                break
        want[c] = val[tuple(src)]
    np.testing.assert_array_equal(got, want)


def test_metrics_and_physics_check():
    from fluxcore3d_amd.solver.thermal_cuda_ext import THERMAL_CUDA
    rng = np.random.default_rng(2)
    n = 100_003
    Tn, To = (rng.standard_normal(n).astype(np.float32) for _ in range(2))
    solid = (rng.random(n) < 0.3).astype(np.uint8)
    rs, rf = THERMAL_CUDA.convergence_metrics(T(Tn), T(To), T(solid))
    d = np.abs(Tn - To)
    assert float(rs.item()) == d[solid == 1].max() and float(rf.item()) == d[solid == 0].max()
    assert int(THERMAL_CUDA.physics_check(T(Tn), -10.0, 10.0).item()) == 0
    bad = Tn.copy()
    bad[77_777] = np.nan
    assert int(THERMAL_CUDA.physics_check(T(bad), -10.0, 10.0).item()) == 1
    assert int(THERMAL_CUDA.physics_check(T(Tn), -10.0, 1.0).item()) == 1
