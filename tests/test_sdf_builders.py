"""Host-side SDF builders of pre_post/sdf_tracing.py:262-341 (CPU)."""
import sys
import types

import numpy as np
import pytest

torch = pytest.importorskip("torch")


def test_phi_from_voxel_mask_is_the_signed_euclidean_distance():
    from fluxcore3d_amd.pre_post.sdf_tracing import build_phi_from_voxel_mask_mm
    rng = np.random.default_rng(3)
    solid = rng.random((6, 5, 7)) < 0.3
    pitch = 0.5
    phi = build_phi_from_voxel_mask_mm(solid, pitch, device="cpu", dtype=torch.float64).numpy()
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in solid.shape], indexing="ij"), -1).reshape(-1, 3)
    s = solid.reshape(-1)
    for n, p in enumerate(idx):
        other = idx[~s] if s[n] else idx[s]
        dmin = np.sqrt(((other - p) ** 2).sum(1)).min() * pitch
        assert abs(phi.reshape(-1)[n] - (-dmin if s[n] else dmin)) < 1e-12


def test_phi_from_stl_samples_the_grid_in_c_order(monkeypatch):
    """The signed distance of the mesh at origin + pitch * (i, j, k), chunked in C
    order; a stand-in for trimesh's ProximityQuery (trimesh is not installed here)
    returns the analytic signed distance of a sphere (positive inside, like trimesh)."""
    from fluxcore3d_amd.pre_post import sdf_tracing
    c, r = np.array([1.0, -0.5, 2.0]), 1.7

    class PQ:
        def __init__(self, mesh):
            assert mesh == "sphere"

        def signed_distance(self, pts):
            return r - np.sqrt(((pts - c) ** 2).sum(1))

    tm = types.ModuleType("trimesh")
    prox = types.ModuleType("trimesh.proximity")
    prox.ProximityQuery = PQ
    tm.proximity = prox
    monkeypatch.setitem(sys.modules, "trimesh", tm)
    monkeypatch.setitem(sys.modules, "trimesh.proximity", prox)
    nx, ny, nz, pitch = 5, 4, 6, 0.4
    org = np.array([0.2, -1.3, 0.9])
    phi = sdf_tracing.build_phi_from_stl_mm("sphere", nx, ny, nz, pitch, org, device="cpu", chunk_size=7).numpy()
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pts = org + pitch * np.stack([i, j, k], -1)
    want = (r - np.sqrt(((pts - c) ** 2).sum(-1))).astype(np.float32)
    np.testing.assert_array_equal(phi, want)
