"""fc_lbm_convective_bc_noalloc: the convective outlet (lbm_kernels.cu:1170-1344
semantics) with its partial sums in the library's static device scratch instead of
a caller workspace. Checked against fc_lbm_convective_bc, the entry point with a
workspace, bit for bit: two calls in a row on the same device, the second on a
smaller grid, so the second reads a scratch array that still holds the first
call's partials."""
import ctypes

import numpy as np
import pytest

from oracle import lbm as olbm

torch = pytest.importorskip("torch")


def _case(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    rho = 1 + 0.01 * rng.random(N)
    u = 0.03 * rng.standard_normal((3, N))
    f = (olbm.feq(rho, *u, np.float64) * (1 + 0.01 * rng.standard_normal((19, N)))).astype(np.float32)
    solid = (rng.random(N) < 0.1).astype(np.uint8)
    return f, solid


def _run(noalloc, f, solid, dims, axis, outlet, interior, cax, rho_out):
    from fluxcore3d_amd import _lib
    nx, ny, nz = dims
    n1, n2 = (ny, nz) if axis == 0 else ((nx, nz) if axis == 1 else (nx, ny))
    ft = torch.as_tensor(f, device="cuda").contiguous()
    st = torch.as_tensor(solid, device="cuda").contiguous()
    stream = _lib.stream_handle(ft.device)
    args = [ft.data_ptr(), st.data_ptr(), nx, ny, nz, axis, outlet, interior, cax, rho_out, n1, n2]
    if noalloc:
        _lib.call("fc_lbm_convective_bc_noalloc", *args, stream)
    else:
        need = int(_lib.call("fc_lbm_convective_bc_workspace_bytes", n1, n2))
        ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        _lib.call("fc_lbm_convective_bc", *args, ws.data_ptr(), ws.numel(), stream)
    torch.cuda.synchronize()
    return ft.cpu().numpy()


@pytest.mark.gpu
def test_noalloc_convective_matches_workspace_entry_twice():
    # first call: 40 x 96 x 80 grid, outlet at x = 39 (a 7680-node plane, 30 blocks)
    # second call: 12 x 20 x 16 grid, outlet at z = 0 (a 240-node plane, 1 block)
    for dims, axis, outlet, interior, cax, seed in (((40, 96, 80), 0, 39, 38, -1, 1), ((12, 20, 16), 2, 0, 1, 1, 2)):
        f, solid = _case(*dims, seed)
        want = _run(False, f, solid, dims, axis, outlet, interior, cax, 1.0)
        got = _run(True, f, solid, dims, axis, outlet, interior, cax, 1.0)
        np.testing.assert_array_equal(got, want)
        assert not np.array_equal(got, f)  # the outlet plane was rewritten
        # and the oracle's convective outlet (tolerance: FP64 partials against a plain FP32 sum)
        ref = olbm.convective(f, solid, *dims, axis, outlet, interior, cax, 1.0)
        assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


def test_noalloc_rejects_planes_beyond_the_scratch_limit():
    """Host-side check, before any device work: a 2048 x 1024 outlet plane exceeds the
    1,048,576-node scratch."""
    from fluxcore3d_amd import _lib
    lib = _lib.lib()
    rc = lib.fc_lbm_convective_bc_noalloc(None, None, 4, 2048, 1024, 0, 3, 2, -1, ctypes.c_float(1.0), 2048, 1024,
                                          None)
    assert rc != 0
    assert b"scratch limit" in lib.fc_last_error()
