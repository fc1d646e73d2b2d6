"""The condition behind the face form of the multigrid coefficients, on the
oracle's restatement of the reference's coefficient build
(nlevel_mg_cuda.py:418-507): with uniform rho*c_p and no advection, the west
coefficient of every cell equals, bit for bit, the east coefficient of its -x
neighbour (likewise for y and z), and the mirrored boundary ghosts reuse the face
towards the mirrored neighbour. So one array per axis holds all six, with the
last plane's slot holding the self-ghost low face of the flow axis. A varying
rho*c_p or any advection breaks it; the library then keeps the per-cell arrays
(it checks every cell on the GPU)."""
import numpy as np
import pytest

from oracle import mg


def _mismatches(k, rcp, u, ax, out_idx):
    z = np.zeros_like(k)
    co = mg.build_coeffs_full(k, rcp, u, z, z, ax, out_idx, 1.0, 0.7)
    per_cell = dict(zip(("E", "W", "N", "S", "U", "D"), co[:6]))
    bad = 0
    for dim, (hi, lo) in enumerate((("E", "W"), ("N", "S"), ("U", "D"))):
        n = k.shape[dim]
        mirror = ax != dim or out_idx == 0  # low ghost = index 1, else the cell itself
        a_hi = np.moveaxis(per_cell[hi], dim, 0)
        a_lo = np.moveaxis(per_cell[lo], dim, 0)
        face = a_hi.copy()  # face between c and c+1 along dim
        face[n - 1] = 0.0 if mirror else a_lo[0]
        got_hi = np.concatenate([face[:n - 1], face[n - 2:n - 1]])  # high ghost mirrors to n-2
        got_lo = np.concatenate([face[0:1] if mirror else face[n - 1:n], face[:n - 1]])
        bad += int(np.sum(got_hi.view(np.uint32) != a_hi.view(np.uint32)))
        bad += int(np.sum(got_lo.view(np.uint32) != a_lo.view(np.uint32)))
    diag = ((((co[0] + co[1]) + co[2]) + co[3]) + co[4]) + co[5]
    bad += int(np.sum(diag.view(np.uint32) != co[6].view(np.uint32)))
    return bad


@pytest.mark.parametrize("ax", [0, 1, 2])
@pytest.mark.parametrize("out_idx_hi", [False, True])
def test_face_form_exact_iff_symmetric(ax, out_idx_hi):
    rng = np.random.default_rng(ax)
    shape = (12, 10, 14)
    k = (10.0 ** rng.uniform(0, 4, shape)).astype(np.float32)  # contrast up to 1e4
    out_idx = shape[ax] - 1 if out_idx_hi else 0
    zero = np.zeros_like(k)
    assert _mismatches(k, np.ones_like(k), zero, ax, out_idx) == 0
    assert _mismatches(k, np.full_like(k, 3.0), zero, ax, out_idx) == 0
    assert _mismatches(k, (1 + rng.random(shape)).astype(np.float32), zero, ax, out_idx) > 0
    assert _mismatches(k, np.ones_like(k), rng.standard_normal(shape).astype(np.float32), ax, out_idx) > 0
