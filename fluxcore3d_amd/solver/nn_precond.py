"""Learned preconditioner for the conduction system A x = b: a U-Net on
channels-last BF16 activations, BF16 weights and FP32 accumulation, run by the
MFMA implicit-GEMM convolutions of csrc/conv.hip, used inside flexible CG with
FP64 outer vectors (mg_pcg.fcg_fp64) and a GMG-PCG fallback.

There is no counterpart in the reference: its thermal solve stops at the
multigrid-preconditioned Krylov methods (solver/mg_pcg.py:33-187). The
network works on the symmetrically Jacobi-scaled system (D = diag(A)):
with r~ = D^-1/2 r and s = ||r~|| / sqrt(n) it returns z = D^-1/2 s N(r~ / s, log k, mask)
(the residual channel at unit RMS, like the O(1) feature channels),
an approximation of A^-1 r, so one application costs a fixed number of
convolutions regardless of the conductivity contrast.

Architecture (levels l = 0..L-1, channels C_l, 2x average pooling between
levels, nearest upsampling + skip concatenation on the way up, leaky ReLU
with slope 0.1, no biases):

  enc_l:  conv(C_{l-1} or 8 -> C_l), conv(C_l -> C_l)
  dec_l:  conv(C_{l+1} + C_l -> C_l), conv(C_l -> C_l)      (l < L-1)
  head:   conv(C_0 -> 8), linear, channel 0 is the output

Training (``train``) runs on the GPU through the same kernels (forward, data
gradient with flipped weights, weight gradient); element-wise work, pooling,
upsampling and concatenation of the training graph are torch operations.
Inference (``UNetPrecond.apply``) uses only HIP kernels: input packing,
convolutions (decoder convolutions read the upsampled coarse tensor and the
skip tensor directly), pooling and output scaling.
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .. import _lib
from . import mg_cuda_ext  # noqa: F401  (registers fc_mg_norm2_f32)

SLOPE = 0.1
IN_CH = 8
HEAD_CH = 8
NN_SLOT = 7  # FP64 scalar slot of the flexible-CG workspace holding ||D^-1/2 r||^2

_lib.register({
    "fc_conv_kpad": (_lib.I32, [_lib.I32]),
    "fc_conv3d_fwd": (_lib.I32, [_lib.P, _lib.I32, _lib.I32, _lib.P, _lib.I32, _lib.P, _lib.I32, _lib.I32,
                                 _lib.I32, _lib.I32, _lib.I32, _lib.I32, _lib.F32, _lib.P, _lib.P, _lib.P]),
    "fc_conv3d_flip_weights": (_lib.I32, [_lib.P, _lib.I32, _lib.I32, _lib.P, _lib.P]),
    "fc_conv3d_dw_workspace_bytes": (_lib.SZ, [_lib.I32, _lib.I32, _lib.I32]),
    "fc_conv3d_bwd_weight": (_lib.I32, [_lib.P, _lib.I32, _lib.I32, _lib.P, _lib.I32, _lib.P, _lib.I32, _lib.I32,
                                        _lib.I32, _lib.I32, _lib.I32, _lib.I32, _lib.P, _lib.P, _lib.SZ, _lib.P]),
    "fc_unet_pack_input": (_lib.I32, [_lib.P, _lib.P, _lib.P, _lib.P, _lib.I32, _lib.I64, _lib.P, _lib.P]),
    "fc_unet_avgpool2": (_lib.I32, [_lib.P, _lib.I32, _lib.I32, _lib.I32, _lib.I32, _lib.I32, _lib.P, _lib.P]),
    "fc_unet_unpack": (_lib.I32, [_lib.P, _lib.I32, _lib.P, _lib.P, _lib.P, _lib.I32, _lib.I64, _lib.P, _lib.P]),
    "fc_unet_scale": (_lib.I32, [_lib.P, _lib.P, _lib.I64, _lib.P, _lib.P]),
    "fc_unet_unpack_add": (_lib.I32, [_lib.P, _lib.I32, _lib.P, _lib.P, _lib.P, _lib.I32, _lib.I64, _lib.P, _lib.P]),
    "fc_unet_residual_scale": (_lib.I32, [_lib.P, _lib.P, _lib.P, _lib.P, _lib.I64, _lib.P, _lib.P]),
}, unchecked={"fc_conv_kpad", "fc_conv3d_dw_workspace_bytes"})


def _stream(dev) -> int:
    return torch.cuda.current_stream(dev).cuda_stream


def _ptr(t: Optional[torch.Tensor]) -> int:
    return 0 if t is None else t.data_ptr()


def kpad(cin: int) -> int:
    return int(_lib.call("fc_conv_kpad", cin))


def pack_weights(w: torch.Tensor) -> torch.Tensor:
    """[cout, 27, cin] FP32 -> BF16 [cout, kpad(cin)] with zero padding."""
    cout, _, cin = w.shape
    out = torch.zeros(cout, kpad(cin), device=w.device, dtype=torch.bfloat16)
    out[:, :27 * cin] = w.detach().reshape(cout, 27 * cin).to(torch.bfloat16)
    return out


def conv3d(x0: torch.Tensor, wpk: torch.Tensor, cout: int, act: bool, x1: Optional[torch.Tensor] = None,
           up0: bool = False, out_f32: bool = False, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """3x3x3 convolution, zero padding. x0 [B, D', H', W', c0] BF16 (half-size
    when up0), x1 optional skip [B, D, H, W, c1] BF16; returns [B, D, H, W, cout]."""
    if x0.dtype != torch.bfloat16 or not x0.is_contiguous():
        raise ValueError("conv3d: x0 must be contiguous bfloat16")
    B = x0.shape[0]
    D, H, W = x0.shape[1:4]
    if up0:
        D, H, W = 2 * D, 2 * H, 2 * W
    c0 = x0.shape[4]
    c1 = 0
    if x1 is not None:
        if x1.dtype != torch.bfloat16 or not x1.is_contiguous() or tuple(x1.shape[:4]) != (B, D, H, W):
            raise ValueError("conv3d: x1 must be contiguous bfloat16 of the output extents")
        c1 = x1.shape[4]
    if wpk.shape != (cout, kpad(c0 + c1)):
        raise ValueError(f"conv3d: packed weights {tuple(wpk.shape)} do not match ({cout}, {kpad(c0 + c1)})")
    if out is None:
        out = torch.empty((B, D, H, W, cout), device=x0.device, dtype=torch.float32 if out_f32 else torch.bfloat16)
    f32 = out.dtype == torch.float32
    _lib.call("fc_conv3d_fwd", x0.data_ptr(), c0, int(up0), _ptr(x1), c1, wpk.data_ptr(), cout, B, D, H, W,
              int(act), SLOPE, 0 if f32 else out.data_ptr(), out.data_ptr() if f32 else 0, _stream(x0.device))
    return out


def conv3d_weight_grad(x0: torch.Tensor, dy16: torch.Tensor, cout: int, n_chunks: int = 0) -> torch.Tensor:
    """dW [cout, 27, cin] (FP32) for a single-source convolution."""
    B, D, H, W, cin = x0.shape
    if n_chunks <= 0:
        n_chunks = max(1, min(256, (B * D * H) // 8))
    ws_bytes = int(_lib.call("fc_conv3d_dw_workspace_bytes", cout, cin, n_chunks))
    ws = torch.empty(ws_bytes // 4 + 1, device=x0.device, dtype=torch.float32)
    dw = torch.empty((cout, 27, cin), device=x0.device, dtype=torch.float32)
    _lib.call("fc_conv3d_bwd_weight", x0.data_ptr(), cin, 0, 0, 0, dy16.data_ptr(), cout, B, D, H, W, n_chunks,
              dw.data_ptr(), ws.data_ptr(), ws_bytes, _stream(x0.device))
    return dw


def conv3d_data_grad(dy16: torch.Tensor, wpk: torch.Tensor, cin: int) -> torch.Tensor:
    """dX [B, D, H, W, cin] (FP32): the forward kernel on flipped, transposed weights."""
    cout = dy16.shape[4]
    wf = torch.empty((cin, kpad(cout)), device=dy16.device, dtype=torch.bfloat16)
    _lib.call("fc_conv3d_flip_weights", wpk.data_ptr(), cout, cin, wf.data_ptr(), _stream(dy16.device))
    return conv3d(dy16, wf, cin, act=False, out_f32=True)


class _ConvFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, act, out_f32, need_dx):
        cout, _, cin = w.shape
        wpk = pack_weights(w)
        y = conv3d(x, wpk, cout, act, out_f32=out_f32)
        ctx.save_for_backward(x, wpk, y if act else None)
        ctx.act, ctx.cout, ctx.cin, ctx.need_dx = act, cout, cin, need_dx
        return y

    @staticmethod
    def backward(ctx, gy):
        x, wpk, y = ctx.saved_tensors
        g = gy.float()
        if ctx.act:
            g = torch.where(y.float() > 0, g, g * SLOPE)
        g16 = g.to(torch.bfloat16).contiguous()
        dw = conv3d_weight_grad(x, g16, ctx.cout)
        dx = conv3d_data_grad(g16, wpk, ctx.cin).to(torch.bfloat16) if ctx.need_dx else None
        return dx, dw, None, None, None


def _pool(x: torch.Tensor) -> torch.Tensor:
    B, D, H, W, C = x.shape
    return x.float().view(B, D // 2, 2, H // 2, 2, W // 2, 2, C).mean((2, 4, 6)).to(torch.bfloat16)


def _up(x: torch.Tensor) -> torch.Tensor:
    B, D, H, W, C = x.shape
    return x[:, :, None, :, None, :, None, :].expand(B, D, 2, H, 2, W, 2, C).reshape(B, 2 * D, 2 * H, 2 * W, C)


CONFIG5_CHANNELS = (16, 16, 24, 32)  # 4 levels, 16 base channels: 198 720 weights


class UNetPrecond:
    """U-Net preconditioner. ``channels`` per level (multiples of 8, <= 64); the
    default is the 4-level, 16-base-channel network of about 200 k weights."""

    def __init__(self, channels: Sequence[int] = CONFIG5_CHANNELS, device="cuda", seed: int = 0):
        self.channels = [int(c) for c in channels]
        for c in self.channels:
            if c % 8 or c > 64:
                raise ValueError("channels must be multiples of 8 and at most 64")
        self.device = torch.device(device)
        self.levels = len(self.channels)
        self.layers = self._layout()
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.params: List[torch.Tensor] = []
        for name, cin, cout, act in self.layers:
            std = math.sqrt(2.0 / (27 * cin)) if act else math.sqrt(1.0 / (27 * cin))
            w = torch.randn((cout, 27, cin), generator=g) * std
            self.params.append(w.to(self.device).requires_grad_(True))
        self.feat_scale = 0.25  # log10(k) scaling of the second input channel
        self.out_gain = 1.0  # output scale, folded into the packed head weights
        self._packed: Optional[List[torch.Tensor]] = None
        self._bufs: Dict[Tuple[int, int, int], dict] = {}

    def _layout(self):
        C = self.channels
        layers = []
        for l in range(self.levels):
            cin = IN_CH if l == 0 else C[l - 1]
            layers.append((f"enc{l}a", cin, C[l], True))
            layers.append((f"enc{l}b", C[l], C[l], True))
        for l in range(self.levels - 2, -1, -1):
            layers.append((f"dec{l}a", C[l + 1] + C[l], C[l], True))
            layers.append((f"dec{l}b", C[l], C[l], True))
        layers.append(("head", C[0], HEAD_CH, False))
        return layers

    def n_params(self) -> int:
        return int(sum(p.numel() for p in self.params))

    def flops(self, shape) -> float:
        """Useful forward FLOPs of one application on a D x H x W grid: sum over
        layers of 2 * 27 * c_in * c_out per voxel of the layer's level."""
        D, H, W = (int(v) for v in shape)
        total = 0.0
        for name, cin, cout, _ in self.layers:
            lvl = int(name[3]) if name[:3] in ("enc", "dec") else 0
            total += 2.0 * 27 * cin * cout * (D >> lvl) * (H >> lvl) * (W >> lvl)
        return total

    # ------------------------------------------------------------ training graph
    def forward_train(self, x: torch.Tensor) -> torch.Tensor:
        """x [B, D, H, W, 8] BF16 -> [B, D, H, W, 8] FP32 (channel 0 is the output)."""
        p = iter(self.params)
        skips = []
        h = x
        for l in range(self.levels):
            if l > 0:
                h = _pool(h)
            h = _ConvFn.apply(h, next(p), True, False, l > 0)
            h = _ConvFn.apply(h, next(p), True, False, True)
            skips.append(h)
        for l in range(self.levels - 2, -1, -1):
            h = torch.cat([_up(h), skips[l]], dim=-1).contiguous()
            h = _ConvFn.apply(h, next(p), True, False, True)
            h = _ConvFn.apply(h, next(p), True, False, True)
        return _ConvFn.apply(h, next(p), False, True, True) * self.out_gain

    # ------------------------------------------------------------------ inference
    def refresh(self) -> None:
        """Re-pack the BF16 weights after a parameter update."""
        self._packed = [pack_weights(w) for w in self.params[:-1]] + [pack_weights(self.params[-1] * self.out_gain)]

    def _buffers(self, shape) -> dict:
        key = tuple(int(s) for s in shape)
        if key not in self._bufs:
            D, H, W = key
            f = 2 ** (self.levels - 1)
            if D % f or H % f or W % f:
                raise ValueError(f"grid {key} must be divisible by {f} for {self.levels} levels")
            bf = dict(device=self.device, dtype=torch.bfloat16)
            b = {"x": torch.empty((1, D, H, W, IN_CH), **bf), "enc": [], "tmp": [], "pool": [], "dec": []}
            for l, c in enumerate(self.channels):
                s = (1, D >> l, H >> l, W >> l)
                b["tmp"].append(torch.empty(s + (c,), **bf))
                b["enc"].append(torch.empty(s + (c,), **bf))
                b["dec"].append(torch.empty(s + (c,), **bf))
                if l > 0:
                    b["pool"].append(torch.empty(s + (self.channels[l - 1],), **bf))
            b["head"] = torch.empty((1, D, H, W, HEAD_CH), device=self.device, dtype=torch.float32)
            b["rs"] = torch.empty(D * H * W, device=self.device, dtype=torch.float32)
            self._bufs[key] = b
        return self._bufs[key]

    def apply(self, r32: torch.Tensor, z32: torch.Tensor, feat: Optional[torch.Tensor], mask: Optional[torch.Tensor],
              dsq: torch.Tensor, scal: torch.Tensor, partial: torch.Tensor, shape,
              Lo: Optional[torch.Tensor] = None) -> None:
        """z = D^-1/2 s N(r~ / s, feat, mask) * mask with r~ = D^-1/2 r (dsq = diag(A)^-1/2)
        and s = ||r~|| kept in scal[NN_SLOT] (FP64, device); ``partial`` is the
        block-partial buffer of the FP64 reductions. With ``Lo`` = A z (z a V-cycle's
        output already in z32) the network sees the residual r - Lo that z leaves
        and its correction is added to z32."""
        if self._packed is None:
            self.refresh()
        D, H, W = shape
        n = D * H * W
        b = self._buffers(shape)
        st = _stream(self.device)
        if Lo is None:
            _lib.call("fc_unet_scale", r32.data_ptr(), dsq.data_ptr(), n, b["rs"].data_ptr(), st)
        else:
            _lib.call("fc_unet_residual_scale", r32.data_ptr(), Lo.data_ptr(), _ptr(mask), dsq.data_ptr(), n,
                      b["rs"].data_ptr(), st)
        _lib.call("fc_mg_norm2_f32", b["rs"].data_ptr(), n, partial.data_ptr(), scal.data_ptr(), NN_SLOT, st)
        _lib.call("fc_unet_pack_input", b["rs"].data_ptr(), _ptr(feat), _ptr(mask), scal.data_ptr(), NN_SLOT, n,
                  b["x"].data_ptr(), st)
        wp = iter(self._packed)
        C = self.channels
        h = b["x"]
        for l in range(self.levels):
            if l > 0:
                src = b["enc"][l - 1]
                _lib.call("fc_unet_avgpool2", src.data_ptr(), 1, D >> (l - 1), H >> (l - 1), W >> (l - 1), C[l - 1],
                          b["pool"][l - 1].data_ptr(), st)
                h = b["pool"][l - 1]
            conv3d(h, next(wp), C[l], True, out=b["tmp"][l])
            conv3d(b["tmp"][l], next(wp), C[l], True, out=b["enc"][l])
        h = b["enc"][self.levels - 1]
        for l in range(self.levels - 2, -1, -1):
            conv3d(h, next(wp), C[l], True, x1=b["enc"][l], up0=True, out=b["tmp"][l])
            conv3d(b["tmp"][l], next(wp), C[l], True, out=b["dec"][l])
            h = b["dec"][l]
        conv3d(h, next(wp), HEAD_CH, False, out=b["head"])
        _lib.call("fc_unet_unpack" if Lo is None else "fc_unet_unpack_add", b["head"].data_ptr(), HEAD_CH,
                  _ptr(mask), dsq.data_ptr(), scal.data_ptr(), NN_SLOT, n, z32.data_ptr(), st)

    # --------------------------------------------------------------- persistence
    def state(self) -> Dict[str, np.ndarray]:
        out = {"channels": np.asarray(self.channels, np.int32), "feat_scale": np.asarray([self.feat_scale]),
               "out_gain": np.asarray([self.out_gain])}
        for (name, *_), w in zip(self.layers, self.params):
            out[name] = w.detach().float().cpu().numpy().astype(np.float16)  # inference runs BF16 weights
        return out

    def save(self, path: str) -> None:
        np.savez_compressed(path, **self.state())

    @classmethod
    def load(cls, path: str, device="cuda") -> "UNetPrecond":
        with np.load(path, allow_pickle=False) as z:
            net = cls(tuple(int(c) for c in z["channels"]), device=device)
            net.feat_scale = float(z["feat_scale"][0])
            net.out_gain = float(z["out_gain"][0]) if "out_gain" in z.files else 1.0
            for (name, *_), w in zip(net.layers, net.params):
                if tuple(z[name].shape) != tuple(w.shape):
                    raise ValueError(f"{path}: layer {name} has shape {z[name].shape}, expected {tuple(w.shape)}")
                with torch.no_grad():
                    w.copy_(torch.from_numpy(z[name].astype(np.float32)).to(device))
        net.refresh()
        return net


def jacobi_scale(mg) -> torch.Tensor:
    """diag(A)^-1/2 of the finest level, A = I + dt L (masked rows: 1)."""
    lv0 = mg._levels[0]
    d = 1.0 + float(mg._dt_cache) * lv0["diagL"].reshape(-1).double()
    m = lv0["mask_f"].reshape(-1) > 0
    return torch.where(m, d.clamp_min(1e-30).rsqrt(), torch.ones_like(d)).float().contiguous()


def log_k_feature(k: torch.Tensor, scale: float = 0.25) -> torch.Tensor:
    """Second input channel: scale * log10(k), FP32, flattened."""
    return (torch.log10(k.float().clamp_min(1e-30)) * scale).reshape(-1).contiguous()


# ------------------------------------------------------------------ training data
def _mg_for(c, device):
    from .nlevel_mg_cuda import NLevelGeometricMGSolver

    mg = NLevelGeometricMGSolver(device, torch.float32, c["ax"], c["sign"], c["s"], omega=0.8, n_pre=2, n_post=2,
                                 n_coarse=20)
    mg.build_if_needed(c["k"], c["rcp"], c["u"], c["v"], c["w"], c["mask"], c["dt"])
    return mg


class ScaledOperator:
    """A~ u = D^-1/2 A D^-1/2 u on the finest level, written with torch tensor
    operations for the training loss only (same stencil, ghosts and masking as
    the HIP operator; the solver never calls this)."""

    def __init__(self, mg):
        lv0 = mg._levels[0]
        self.shape = tuple(int(v) for v in lv0["shape"])
        self.c = {k: lv0[k].reshape(self.shape).float().clone() for k in ("aE", "aW", "aN", "aS", "aU", "aD",
                                                                           "diagL")}
        self.mask = lv0["mask_f"].reshape(self.shape).float().clone()
        self.dt = float(mg._dt_cache)
        self.dsq = jacobi_scale(mg).reshape(self.shape).clone()
        dev = self.mask.device
        self.idx = []
        for ax, n in enumerate(self.shape):
            lo = 1 if (mg.ax != ax or int(lv0["out_idx"]) == 0) else 0
            up = torch.cat([torch.arange(1, n), torch.tensor([n - 2])]).to(dev)
            dn = torch.cat([torch.tensor([lo]), torch.arange(0, n - 1)]).to(dev)
            self.idx.append((up, dn))

    def __call__(self, u: torch.Tensor) -> torch.Tensor:
        v = u * self.dsq
        c = self.c
        nb = []
        for ax in range(3):
            up, dn = self.idx[ax]
            nb.append((v.index_select(v.dim() - 3 + ax, up), v.index_select(v.dim() - 3 + ax, dn)))
        L = (c["diagL"] * v - c["aE"] * nb[0][0] - c["aW"] * nb[0][1] - c["aN"] * nb[1][0] - c["aS"] * nb[1][1]
             - c["aU"] * nb[2][0] - c["aD"] * nb[2][1])
        return (v + self.dt * L) * self.mask * self.dsq


def harvest_pairs(n: int = 128, problems: int = 6, seed0: int = 1000, per_problem: int = 24, device="cuda",
                  feat_scale: float = 0.25, log=print):
    """(input, target, mask) samples at n^3 on smooth-noise blob problems with
    k_s / k_f log-uniform in [1, 1e4]: random errors e (white noise filtered by
    a Gaussian of width sigma, sigma = 0 or log-uniform in [0.5, 32] voxels,
    masked) with r = A e (FP64 operator of the finest level).

    The errors are independent of the problem's source: residual / error
    snapshots of a solve are not used, because their errors are predictable
    from the geometry channels alone and the network then ignores the residual.

    In the Jacobi-scaled variables r~ = D^-1/2 r, e~ = D^1/2 e, inputs are
    [n, n, n, 8] BF16 (r~ at unit RMS, feat_scale log10 k, mask); targets are e~
    with the same scale. Returns (samples, operators)."""
    from .. import geometry

    samples = []
    ops = []
    gen = torch.Generator(device="cpu").manual_seed(seed0)
    for i in range(problems):
        c = geometry.conduction_problem(n, seed=seed0 + i, device=device)
        mg = _mg_for(c, device)
        lv0 = mg._levels[0]
        mask = lv0["mask_f"].reshape(-1)
        feat = log_k_feature(c["k"], feat_scale)
        dsq = jacobi_scale(mg).double()
        ops.append(ScaledOperator(mg))
        pid = len(ops) - 1
        for _ in range(per_problem):
            e = _smooth_noise(n, gen, device).reshape(-1).double() * mask.double()
            r = apply_A64(mg, e)
            samples.append(_sample(r * dsq, e / dsq, feat, mask, n, pid))
        if log:
            log(f"[harvest] problem {i}: k ratio {c['ratio']:.3g}, {len(samples)} samples")
        del mg
    return samples, ops


def _sample(r64, e64, feat, mask, n, pid):
    """(input BF16 [n,n,n,8], target e~ sqrt(N)/||r~|| FP32, mask, problem id, e~.r~/||r~||^2): the
    residual channel at unit RMS, as the inference packing kernel makes it."""
    rn = float(r64.norm()) / float(r64.numel()) ** 0.5
    x = torch.zeros((n, n, n, IN_CH), device=r64.device, dtype=torch.float32)
    x[..., 0] = (r64 / rn).float().view(n, n, n)
    x[..., 1] = feat.view(n, n, n)
    x[..., 2] = mask.view(n, n, n)
    energy = float((e64 * r64).sum()) / rn ** 2
    return x.to(torch.bfloat16), (e64 / rn).float().view(n, n, n), mask.view(n, n, n), pid, energy


def _smooth_noise(n, gen, device):
    """Unit-RMS Gaussian-filtered white noise; sigma = 0 (white) with probability
    1/8, else log-uniform in [0.5, 32] voxels."""
    noise = torch.randn((n, n, n), generator=gen).to(device)
    if float(torch.rand(1, generator=gen)) < 0.125:
        return noise
    sigma = float(2.0 ** torch.empty(1).uniform_(-1.0, 5.0, generator=gen))
    k = torch.fft.fftfreq(n, device=device)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    f = torch.fft.ifftn(torch.fft.fftn(noise) * torch.exp(-2 * (math.pi * sigma) ** 2 * k2)).real
    return f / (f.std() + 1e-30)


def apply_A64(mg, e64: torch.Tensor) -> torch.Tensor:
    """A e on the finest level (FP64 vectors, FP32 coefficients), masked."""
    from .mg_pcg import FCGWorkspace

    lv0 = mg._levels[0]
    nx, ny, nz = lv0["shape"]
    ws = FCGWorkspace(lv0["shape"], mg.device)
    out = torch.empty_like(e64)
    _lib.call("fc_mg_apply_A64", e64.data_ptr(), out.data_ptr(), 0, lv0["mask_f"].data_ptr(), lv0["coef"],
              float(mg._dt_cache), nx, ny, nz, mg.ax, lv0["out_idx"], ws.partial.data_ptr(), ws.scal.data_ptr(),
              _stream(mg.device))
    return out


def train(net: UNetPrecond, samples, ops, steps: int = 600, batch: int = 2, lr: float = 3e-4, seed: int = 0,
          log_every: int = 50, log=print, clip: float = 1.0):
    """Adam on the one-step residual reduction of the network's direction:
    for z = N(r~), the best multiple of A~ z removes the part of r~ along it, so

        loss = 1 - (r~ . A~ z)^2 / (|r~|^2 |A~ z|^2)  in [0, 1],

    the squared sine between r~ and A~ z. It is invariant to the scale of z (as
    flexible CG is), so the large and sample-dependent dynamic range of A~^-1
    does not enter. (An energy norm is not available: the operator is not
    symmetric and its symmetric part is indefinite near Dirichlet and reflected
    Neumann rows.) The output gain is set afterwards to the least-squares
    optimum, which only matters for callers that use z without a line search."""
    opt = torch.optim.Adam(net.params, lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=steps, eta_min=lr * 0.05)
    gen = torch.Generator(device="cpu").manual_seed(seed)
    net.out_gain = 1.0
    hist = []
    for step in range(steps):
        idx = torch.randint(len(samples), (batch,), generator=gen).tolist()
        x = torch.stack([samples[i][0] for i in idx]).contiguous()
        y = net.forward_train(x)[..., 0]
        loss = 0.0
        for b, i in enumerate(idx):
            _, e, m, pid, _ = samples[i]
            Az = ops[pid](y[b] * m)
            with torch.no_grad():
                r = ops[pid](e * m)
            c = (Az * r).sum()
            loss = loss + (1.0 - c * c / ((Az * Az).sum() * (r * r).sum() + 1e-30)) / batch
        opt.zero_grad(set_to_none=True)
        loss.backward()
        torch.nn.utils.clip_grad_norm_(net.params, clip)
        opt.step()
        sched.step()
        hist.append(float(loss.detach()))
        if log and (step % log_every == 0 or step == steps - 1):
            log(f"[train] step {step} loss {hist[-1]:.4f}")
    with torch.no_grad():
        gains = []
        for x, e, m, pid, _ in samples[:: max(1, len(samples) // 8)]:
            Az = ops[pid](net.forward_train(x[None].contiguous())[0, ..., 0] * m)
            r = ops[pid](e * m)
            gains.append(float((Az * r).sum() / ((Az * Az).sum() + 1e-30)))
        net.out_gain = float(np.median(gains))
    net.refresh()
    return hist


# ---------------------------------------------------------------- solver driver
def nn_fcg(mg, net: UNetPrecond, b: torch.Tensor, k: torch.Tensor, tol: float = 1e-8, max_it: int = 300,
           x0: Optional[torch.Tensor] = None, stall_window: Optional[int] = None, stall_factor: float = 0.5,
           m: int = 6, ws=None, check_every: int = 1, hybrid: bool = True):
    """Flexible GCR(m) with the U-Net preconditioner (FP64 vectors; the FP64 true
    residual decides convergence). The operator is not symmetric and a learned
    preconditioner is not exact, so a minimal-residual flexible method is used
    (flexible CG does not converge with weaker preconditioners on this
    operator): every step removes the part of r along the orthogonalised
    A-image of the network's direction. With ``hybrid`` (the default, the
    network trained by train_gmg) each direction is one V-cycle plus the network's
    correction of the residual the V-cycle leaves (hybrid_precond); without it the
    network alone (trained by train). ``stall_window`` defaults to 8 iterations for
    the network alone and 24 for the hybrid, whose first iterations plateau like
    GMG-GCR's own (at 512^3 the residual stays near 1 for about six). When the residual
    has not dropped by ``stall_factor`` over
    ``stall_window`` iterations, or the budget ends unconverged, the solve
    continues from the current iterate (never worse than the start) with the
    V-cycle as preconditioner. Returns (x, info) with info["fallback"]."""
    from .mg_pcg import GCRWorkspace, gcr_fp64

    lv0 = mg._levels[0]
    shape = lv0["shape"]
    if stall_window is None:
        stall_window = 24 if hybrid else 8
    if ws is None:
        ws = GCRWorkspace(shape, mg.device, m)
    feat = log_k_feature(k, net.feat_scale)
    mask = lv0["mask_f"].reshape(-1)
    dsq = jacobi_scale(mg)
    hist_it = []
    nn_scal = torch.zeros(8, dtype=torch.float64, device=mg.device)

    if hybrid:
        precond = hybrid_precond(mg, net, k)
    else:
        def precond(r32, z32, w):
            net.apply(r32, z32, feat, mask, dsq, nn_scal, w.fcg.partial, shape)

    def watch(it, rel):
        hist_it.append(rel)
        if len(hist_it) > stall_window and hist_it[-1] > stall_factor * hist_it[-1 - stall_window]:
            raise StopIteration

    x, info = gcr_fp64(mg, b, x0=x0, tol=tol, max_it=max_it, precond=precond, m=m, ws=ws, on_iter=watch,
                       check_every=check_every)
    it_nn = info["it"]
    if info["converged"]:
        info["fallback"] = False
        info["it_nn"] = it_nn
        return x, info
    x2, info2 = gcr_fp64(mg, b, x0=ws.fcg.x.view(shape).clone(), tol=tol, max_it=max_it, m=m)
    info2["fallback"] = True
    info2["it_nn"] = it_nn
    info2["stalled"] = info.get("stopped", False)
    info2["res_nn"] = info["res_rel"]
    return x2, info2


# ------------------------------------------------- training on GMG-generated residuals
def harvest_gmg(n: int = 128, seed: int = 0, device="cuda", max_it: int = 400, tol: float = 1e-10,
                stride: int = 1, feat_scale: float = 0.25, pid: int = 0):
    """Training samples generated by the GMG solver itself: one smooth-noise blob
    conduction problem at n^3 (k_s / k_f log-uniform in [1, 1e4] from ``seed``) is
    solved with GMG-preconditioned flexible CG, and at every ``stride``-th iteration
    the residual r_k is passed through one V-cycle, z = V(r_k), leaving the
    post-V-cycle residual r' = r_k - A z (the part the V-cycle misses, which the
    network learns to remove: A e' = r' for the remaining error e'). Each sample is
    (input [n, n, n, 8] BF16 = {r~' at unit RMS, feat_scale log10 k, mask}, target
    r~' at unit RMS FP32, mask, problem id) with r~' = D^-1/2 r'. Returns
    (samples, ScaledOperator of the problem, info)."""
    from .. import geometry
    from .mg_pcg import fcg_fp64

    c = geometry.conduction_problem(n, seed=seed, device=device)
    mg = _mg_for(c, device)
    lv0 = mg._levels[0]
    shape = lv0["shape"]
    N = n ** 3
    mask = lv0["mask_f"].reshape(-1)
    feat = log_k_feature(c["k"], feat_scale)
    dsq = jacobi_scale(mg)
    z = torch.empty(N, device=device)
    Lo = torch.empty(N, device=device)
    rs = torch.empty(N, device=device)
    st = _stream(mg.device)
    samples = []

    def grab(it, ws):
        if it % stride:
            return
        mg.precondition(ws.r32, out=z, masked_input=True)
        mg._Mop_cuda(z.view(shape), 0, Lo.view(shape))
        _lib.call("fc_unet_residual_scale", ws.r32.data_ptr(), Lo.data_ptr(), mask.data_ptr(), dsq.data_ptr(), N,
                  rs.data_ptr(), st)
        rms = float(rs.double().norm()) / N ** 0.5
        if not rms > 0.0:
            return
        t = (rs / rms).view(n, n, n)
        x = torch.zeros((n, n, n, IN_CH), device=device, dtype=torch.bfloat16)
        x[..., 0] = t
        x[..., 1] = feat.view(n, n, n)
        x[..., 2] = mask.view(n, n, n)
        samples.append((x, t.clone(), mask.view(n, n, n), pid))

    _, info = fcg_fp64(mg, c["b"], tol=tol, max_it=max_it, on_iter=grab)
    op = ScaledOperator(mg)
    del mg
    return samples, op, {"ratio": float(c["ratio"]), "gmg_it": int(info["it"]), "samples": len(samples)}


def train_gmg(net: UNetPrecond, n: int = 128, problems: int = 32, seed0: int = 1000, steps_per_problem: int = 64,
              batch: int = 2, lr: float = 3e-4, replay: int = 256, stride: int = 1, seed: int = 0, log=print,
              clip: float = 1.0):
    """Train on GMG-generated post-V-cycle residuals, problem by problem (each
    problem's samples are made on the fly by harvest_gmg, trained on together with a
    replay reservoir of ``replay`` earlier samples, then dropped). Loss: the residual
    the network's correction leaves, relative to the one it was given,

        loss = |t - A~ N(t)|^2 / |t|^2,   t = r~' at unit RMS,

    so the correction is learned with its scale (it is added to the V-cycle's
    output without a separate line search)."""
    import random

    total = problems * steps_per_problem
    opt = torch.optim.Adam(net.params, lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=total, eta_min=lr * 0.05)
    rng = random.Random(seed)
    net.out_gain = 1.0
    reservoir: list = []
    ops: Dict[int, ScaledOperator] = {}
    hist = []
    n_samples = 0
    for i in range(problems):
        samples, op, info = harvest_gmg(n, seed=seed0 + i, device=net.device, stride=stride, pid=i)
        ops[i] = op
        n_samples += len(samples)
        pool = samples + reservoir
        ploss = []
        for _ in range(steps_per_problem):
            pick = [pool[rng.randrange(len(pool))] for _ in range(batch)]
            x = torch.stack([p[0] for p in pick]).contiguous()
            y = net.forward_train(x)[..., 0]
            loss = 0.0
            for b, (_, t, m, pid) in enumerate(pick):
                res = t - ops[pid](y[b] * m)
                loss = loss + (res * res).sum() / ((t * t).sum() + 1e-30) / batch
            opt.zero_grad(set_to_none=True)
            loss.backward()
            torch.nn.utils.clip_grad_norm_(net.params, clip)
            opt.step()
            sched.step()
            ploss.append(float(loss.detach()))
        hist += ploss
        reservoir = rng.sample(reservoir + samples, min(replay, len(reservoir) + len(samples)))
        keep = {p[3] for p in reservoir}
        for k in [k for k in ops if k not in keep]:
            del ops[k]
        del samples, pool
        if log:
            log(f"[train] problem {i}: k ratio {info['ratio']:.3g}, GMG-PCG {info['gmg_it']} it, "
                f"{info['samples']} samples ({n_samples} so far), loss {np.mean(ploss[:8]):.4f} -> "
                f"{np.mean(ploss[-8:]):.4f}")
    net.refresh()
    return {"loss": hist, "samples": n_samples}


def hybrid_precond(mg, net: UNetPrecond, k: torch.Tensor):
    """z = V(r) + N(r - A V(r)): one V-cycle, then the network's correction of the
    residual the V-cycle leaves (HIP kernels only). Returns precond(r32, z32, ws)
    for gcr_fp64 (ws: its GCRWorkspace)."""
    lv0 = mg._levels[0]
    shape = lv0["shape"]
    feat = log_k_feature(k, net.feat_scale)
    mask = lv0["mask_f"].reshape(-1)
    dsq = jacobi_scale(mg)
    Lo = torch.empty(int(mask.numel()), device=mg.device)
    nn_scal = torch.zeros(8, dtype=torch.float64, device=mg.device)

    def precond(r32, z32, w):
        mg.precondition(r32, out=z32, masked_input=True)
        mg._Mop_cuda(z32.view(shape), 0, Lo.view(shape))
        net.apply(r32, z32, feat, mask, dsq, nn_scal, w.fcg.partial, shape, Lo=Lo)

    return precond
