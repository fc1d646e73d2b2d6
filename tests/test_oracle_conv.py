"""Pins of the convolution oracle: torch.nn.functional.conv3d (library routine),
adjoint identities and finite differences."""
import numpy as np
import pytest

from oracle import conv


def _data(seed=0, B=2, D=5, H=4, W=6, ci=3, co=4):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((B, D, H, W, ci)), rng.standard_normal((co, 27, ci)), rng.standard_normal(co)


def test_forward_matches_torch_conv3d():
    torch = pytest.importorskip("torch")
    x, W, b = _data()
    y = conv.conv_fwd(x, W, b)
    wt = torch.from_numpy(W.reshape(W.shape[0], 3, 3, 3, W.shape[2])).permute(0, 4, 1, 2, 3)
    yt = torch.nn.functional.conv3d(torch.from_numpy(x).permute(0, 4, 1, 2, 3), wt, torch.from_numpy(b), padding=1)
    assert np.abs(y - yt.permute(0, 2, 3, 4, 1).numpy()).max() < 1e-12


def test_backward_is_adjoint_and_matches_finite_differences():
    x, W, b = _data(1)
    rng = np.random.default_rng(2)
    dy = rng.standard_normal(x.shape[:4] + (W.shape[0],))
    # <conv(x), dy> = <x, conv_bwd_data(dy)> (adjoint)
    lhs = float((conv.conv_fwd(x, W) * dy).sum())
    rhs = float((x * conv.conv_bwd_data(dy, W)).sum())
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)
    dW, db = conv.conv_bwd_weight(x, dy)
    # loss = <conv(x; W, b), dy> is linear in W: finite differences are exact up to rounding
    for (o, t, c) in [(0, 0, 0), (3, 13, 2), (1, 26, 1)]:
        Wp = W.copy()
        Wp[o, t, c] += 1e-3
        fd = (float((conv.conv_fwd(x, Wp, b) * dy).sum()) - float((conv.conv_fwd(x, W, b) * dy).sum())) / 1e-3
        assert abs(fd - dW[o, t, c]) < 1e-7 * max(1.0, abs(fd))
    assert np.abs(db - dy.sum(axis=(0, 1, 2, 3))).max() < 1e-12
