"""MFMA convolutions of the learned preconditioner against the FP64 oracle
(oracle/conv.py, itself pinned to torch.nn.functional.conv3d and to adjoint /
finite-difference identities). Inputs and weights are BF16 values; the
kernels accumulate in FP32, so the bound is a few FP32 ulps of the sum of
absolute products: |y - y_ref| <= 1e-5 * (|W| * |x|)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import conv as oconv  # noqa: E402

DEV = "cuda"


def _bf16(shape, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randn(shape, generator=g).to(torch.bfloat16)


def _np(t):
    return t.double().cpu().numpy()


def _bound(x, W):
    return oconv.conv_fwd(np.abs(x), np.abs(W))


def _up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


# W >= 64 runs the LDS-staged kernel (input tile of 3 x 6 x 66 voxels) when the tile
# fits the device's LDS, at most 80 KB: partial h and w tiles, 1-4 channel chunks;
# cin = 40 at W = 96 (95 KB) falls back to the cached kernel. Up to 16 input and 16
# output channels run the four-planes-per-wavefront form: depths that are not a multiple
# of four, a single plane, blocks of 4 and 8 planes.
@pytest.mark.parametrize("B,D,H,W,cin,cout,act", [(2, 6, 5, 19, 8, 12, False), (1, 4, 7, 33, 16, 16, True),
                                                   (1, 3, 3, 16, 24, 48, True), (1, 5, 4, 20, 40, 64, False),
                                                   (1, 3, 2, 70, 16, 24, True), (2, 2, 3, 40, 8, 32, False),
                                                   (1, 5, 6, 64, 8, 8, True), (2, 3, 5, 130, 24, 16, False),
                                                   (1, 2, 9, 96, 40, 64, True), (1, 4, 6, 64, 32, 16, True),
                                                   (1, 3, 5, 130, 32, 8, False), (1, 7, 6, 64, 16, 16, True),
                                                   (2, 9, 5, 130, 16, 12, False), (1, 1, 4, 64, 16, 16, False),
                                                   (1, 10, 3, 70, 8, 16, True), (1, 16, 8, 128, 16, 16, True)])
def test_forward_single_source(B, D, H, W, cin, cout, act):
    from fluxcore3d_amd.solver import nn_precond as nn
    x = _bf16((B, D, H, W, cin), 1)
    w = _bf16((cout, 27, cin), 2).float() * 0.2
    wpk = nn.pack_weights(w.to(DEV))
    y = nn.conv3d(x.to(DEV), wpk, cout, act, out_f32=True).cpu().numpy()
    wr = w.to(torch.bfloat16).double().numpy()
    ref = oconv.conv_fwd(_np(x), wr)
    if act:
        ref = np.where(ref > 0, ref, nn.SLOPE * ref)
    assert np.all(np.abs(y - ref) <= 1e-5 * _bound(_np(x), wr) + 1e-30)
    # BF16 output: the FP32 result rounded once
    y16 = nn.conv3d(x.to(DEV), wpk, cout, act).float().cpu().numpy()
    assert np.all(np.abs(y16 - y) <= 2.0 ** -8 * np.abs(y) + 1e-30)


def test_forward_planes_per_block_against_torch_fp32():
    """A shape large enough for blocks of eight output planes (two steps of four planes
    through the block's ring of input planes): against torch's FP32 conv3d of the same
    BF16 values, to 1e-5 of the convolution of the absolute values."""
    from fluxcore3d_amd.solver import nn_precond as nn
    B, D, H, W, cin, cout = 4, 128, 128, 64, 16, 16
    g = torch.Generator(device="cpu").manual_seed(11)
    x = torch.randn((B, D, H, W, cin), generator=g).to(torch.bfloat16).to(DEV)
    w = (torch.randn((cout, 27, cin), generator=g) * 0.2).to(torch.bfloat16).float().to(DEV)
    y = nn.conv3d(x, nn.pack_weights(w), cout, False, out_f32=True)
    xt = x.float().permute(0, 4, 1, 2, 3).contiguous()
    wt = w.reshape(cout, 3, 3, 3, cin).permute(0, 4, 1, 2, 3).contiguous()
    ref = torch.nn.functional.conv3d(xt, wt, padding=1).permute(0, 2, 3, 4, 1)
    bound = torch.nn.functional.conv3d(xt.abs(), wt.abs(), padding=1).permute(0, 2, 3, 4, 1)
    assert bool(((y - ref).abs() <= 1e-5 * bound + 1e-30).all())


@pytest.mark.parametrize("W,c0,cs,cout", [(16, 16, 8, 24), (64, 16, 8, 24), (64, 16, 16, 24), (128, 16, 16, 24),
                                           (64, 16, 16, 16), (128, 8, 8, 12), (64, 8, 8, 16)])
def test_forward_two_sources_with_upsample(W, c0, cs, cout):
    """Decoder convolution: the half-resolution source read through nearest
    upsampling and the skip source read directly; 16 + 16 channels at W >= 64 is the
    level-0 decoder layer's shape (four channel chunks in the LDS-staged kernel); 8 + 8
    channels into up to 16 outputs run the four-planes-per-wavefront form."""
    from fluxcore3d_amd.solver import nn_precond as nn
    xc = _bf16((1, 4, 3, W // 2, c0), 3)  # half-resolution source
    xs = _bf16((1, 8, 6, W, cs), 4)  # skip source at full resolution
    w = _bf16((cout, 27, c0 + cs), 5).float() * 0.2
    wpk = nn.pack_weights(w.to(DEV))
    y = nn.conv3d(xc.to(DEV), wpk, cout, True, x1=xs.to(DEV), up0=True, out_f32=True).cpu().numpy()
    xcat = np.concatenate([_up(_np(xc)), _np(xs)], axis=-1)
    wr = w.to(torch.bfloat16).double().numpy()
    ref = oconv.conv_fwd(xcat, wr)
    ref = np.where(ref > 0, ref, nn.SLOPE * ref)
    assert np.all(np.abs(y - ref) <= 1e-5 * _bound(xcat, wr) + 1e-30)


@pytest.mark.parametrize("cin,cout", [(8, 16), (24, 8), (56, 24)])
def test_data_gradient(cin, cout):
    from fluxcore3d_amd.solver import nn_precond as nn
    dy = _bf16((2, 5, 4, 18, cout), 6)
    w = _bf16((cout, 27, cin), 7).float() * 0.2
    wpk = nn.pack_weights(w.to(DEV))
    dx = nn.conv3d_data_grad(dy.to(DEV), wpk, cin).cpu().numpy()
    wr = w.to(torch.bfloat16).double().numpy()
    ref = oconv.conv_bwd_data(_np(dy), wr)
    bound = oconv.conv_bwd_data(np.abs(_np(dy)), np.abs(wr))
    assert np.all(np.abs(dx - ref) <= 1e-5 * bound + 1e-30)


@pytest.mark.parametrize("cin,cout", [(8, 16), (16, 12), (40, 24)])
def test_weight_gradient(cin, cout):
    from fluxcore3d_amd.solver import nn_precond as nn
    x = _bf16((2, 6, 5, 37, cin), 8)
    dy = _bf16((2, 6, 5, 37, cout), 9)
    dw = nn.conv3d_weight_grad(x.to(DEV), dy.to(DEV), cout).cpu().numpy()
    ref, _ = oconv.conv_bwd_weight(_np(x), _np(dy))
    bound, _ = oconv.conv_bwd_weight(np.abs(_np(x)), np.abs(_np(dy)))
    assert np.all(np.abs(dw - ref) <= 1e-5 * bound + 1e-30)


def test_training_graph_matches_torch_fp32_reference():
    """U-Net forward and weight gradients of the HIP training graph against the
    same network written with torch.nn.functional.conv3d in FP32 with the same
    BF16 roundings (weights, input, every activation; the casts round the
    activation gradients in the backward pass as the HIP graph does)."""
    from fluxcore3d_amd.solver import nn_precond as nn
    net = nn.UNetPrecond(channels=(8, 16, 16), device=DEV, seed=1)
    x = _bf16((1, 8, 8, 16, 8), 10).to(DEV)
    y = net.forward_train(x)
    loss = (y[..., 0] ** 2).sum()
    loss.backward()
    F = torch.nn.functional

    def conv(h, w, act):
        wt = w.detach().to(torch.bfloat16).float().reshape(w.shape[0], 3, 3, 3, w.shape[2]).permute(0, 4, 1, 2, 3)
        wt = wt.contiguous().requires_grad_(True)
        o = F.conv3d(h.permute(0, 4, 1, 2, 3), wt, padding=1).permute(0, 2, 3, 4, 1)
        if not act:
            return o, wt
        return F.leaky_relu(o, nn.SLOPE).to(torch.bfloat16).float(), wt

    h = x.float()
    skips, wts = [], []
    it = iter(net.params)
    for l in range(net.levels):
        if l > 0:
            B, D, H, W, C = h.shape
            h = h.view(B, D // 2, 2, H // 2, 2, W // 2, 2, C).mean((2, 4, 6)).to(torch.bfloat16).float()
        for _ in range(2):
            h, wt = conv(h, next(it), True)
            wts.append(wt)
        skips.append(h)
    for l in range(net.levels - 2, -1, -1):
        up = h.repeat_interleave(2, 1).repeat_interleave(2, 2).repeat_interleave(2, 3)
        h = torch.cat([up, skips[l]], -1)
        for _ in range(2):
            h, wt = conv(h, next(it), True)
            wts.append(wt)
    yr, wt = conv(h, next(it), False)
    wts.append(wt)
    (yr[..., 0] ** 2).sum().backward()
    rel = (y[..., 0] - yr[..., 0]).norm() / yr[..., 0].norm()
    assert rel < 1e-2, float(rel)
    for p, wt in zip(net.params, wts):
        g_ref = wt.grad.permute(0, 2, 3, 4, 1).reshape(p.shape)
        assert (p.grad - g_ref).norm() <= 2e-2 * g_ref.norm() + 1e-12, (p.shape, float((p.grad - g_ref).norm() / g_ref.norm()))


def test_inference_matches_training_graph():
    """Fused inference (diagonal scaling, FP64 norm on the device, packing,
    two-source decoder convolutions, HIP pooling, output scaling) equals the
    training graph applied to the same scaled input."""
    from fluxcore3d_amd import _lib
    from fluxcore3d_amd.solver import nn_precond as nn
    net = nn.UNetPrecond(channels=(8, 16, 16), device=DEV, seed=2)
    D, H, W = 8, 12, 16
    n = D * H * W
    g = torch.Generator(device="cpu").manual_seed(11)
    r = torch.randn(n, generator=g).to(DEV)
    feat = (torch.rand(n, generator=g) - 0.5).to(DEV)
    mask = (torch.rand(n, generator=g) > 0.1).float().to(DEV)
    dsq = (0.1 + torch.rand(n, generator=g)).to(DEV)
    scal = torch.zeros(8, dtype=torch.float64, device=DEV)
    partial = torch.empty(3 * int(_lib.call("fc_mg_partials_count", n)), dtype=torch.float64, device=DEV)
    z = torch.empty(n, device=DEV)
    net.apply(r, z, feat, mask, dsq, scal, partial, (D, H, W))
    rs = r.double() * dsq.double()
    assert abs(float(scal[nn.NN_SLOT]) ** 0.5 - float(rs.norm())) <= 1e-6 * float(rs.norm())
    s = float(rs.norm()) / n ** 0.5  # the residual enters at unit RMS
    x = torch.zeros((1, D, H, W, 8), device=DEV)
    x[..., 0] = (rs / s).float().view(D, H, W)
    x[..., 1] = feat.view(D, H, W)
    x[..., 2] = mask.view(D, H, W)
    with torch.no_grad():
        y = net.forward_train(x.to(torch.bfloat16))[0, ..., 0].reshape(-1)
    zr = s * y * mask * dsq
    assert (z - zr).norm() <= 1e-2 * zr.norm()
