"""Train the U-Net preconditioner on the GPU and compare NN-FCG with GMG-PCG.

    python -m fluxcore3d_amd.solver.train_unet --out fluxcore3d_amd/data/unet_precond.npz

Default: the 4-level, 16-base-channel network (about 200 k weights) trained on
residuals that the GMG solver generates on the fly at 128^3 (nn_precond.train_gmg):
for each of ``--problems`` smooth-noise blob conduction problems (k_s / k_f
log-uniform in [1, 1e4]) GMG-preconditioned flexible CG runs to 1e-10 and every
iteration's residual, passed through one V-cycle, gives one post-V-cycle residual
sample (about 150 per problem, thousands in all); the network learns the
correction of what the V-cycle leaves. ``--random-errors`` trains the stand-alone
network on random errors instead (nn_precond.harvest_pairs / train).
Evaluation uses problems with seeds outside the training range at 128^3 and
256^3, FP32 V-cycle / FP64 outer vectors, tolerance 1e-8: flexible GCR with the
network against GMG-PCG (flexible CG) and GMG-GCR(6).
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np
import torch


def evaluate(net, n: int, seed: int, device, tol: float = 1e-8, ratio=None, hybrid: bool = True) -> dict:
    from .. import geometry
    from .mg_pcg import FCGWorkspace, fcg_fp64, gcr_fp64
    from .nn_precond import _mg_for, nn_fcg

    c = geometry.conduction_problem(n, seed=seed, ratio=ratio, device=device)
    mg = _mg_for(c, device)
    ws = FCGWorkspace(c["k"].shape, device)
    fcg_fp64(mg, c["b"], tol=tol, max_it=2, ws=ws)
    nn_fcg(mg, net, c["b"], c["k"], tol=tol, max_it=2, hybrid=hybrid)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, gi = fcg_fp64(mg, c["b"], tol=tol, max_it=1000, ws=ws)
    torch.cuda.synchronize()
    t_gmg = time.perf_counter() - t0
    gcr_fp64(mg, c["b"], tol=tol, max_it=2, m=6)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, gg = gcr_fp64(mg, c["b"], tol=tol, max_it=1000, m=6)
    torch.cuda.synchronize()
    t_gcr = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, ni = nn_fcg(mg, net, c["b"], c["k"], tol=tol, max_it=1000, hybrid=hybrid)
    torch.cuda.synchronize()
    t_nn = time.perf_counter() - t0
    return {"n": n, "seed": seed, "ratio": round(float(c["ratio"]), 2), "gmg_s": round(t_gmg, 4),
            "gmg_it": gi["it"], "gmg_res": gi["res_rel"], "gmg_gcr_s": round(t_gcr, 4), "gmg_gcr_it": gg["it"],
            "nn_s": round(t_nn, 4), "nn_it": ni["it"],
            "nn_it_network": ni.get("it_nn"), "nn_fallback": ni["fallback"], "nn_res": ni["res_rel"],
            "nn_converged": bool(ni["converged"])}


def unet_flops(net, n: int) -> float:
    """Forward FLOPs of one application at n^3: sum over layers of 2 * 27 * cin * cout per voxel."""
    total = 0.0
    for name, cin, cout, _ in net.layers:
        lvl = int(name[3]) if name[:3] in ("enc", "dec") else 0
        total += 2.0 * 27 * cin * cout * (n >> lvl) ** 3
    return total


def time_apply(weights: str, n: int, reps: int = 10) -> dict:
    """Time one U-Net application (inference path, all HIP kernels) at n^3."""
    from .nn_precond import UNetPrecond

    dev = torch.device("cuda:0")
    net = UNetPrecond.load(weights, device=dev)
    N = n ** 3
    g = torch.Generator(device="cpu").manual_seed(0)
    r = torch.randn(N, generator=g).to(dev)
    z = torch.empty_like(r)
    ones = torch.ones(N, device=dev)
    feat = torch.zeros(N, device=dev)
    scal = torch.zeros(8, dtype=torch.float64, device=dev)
    partial = torch.empty(3 * 8192, dtype=torch.float64, device=dev)
    for _ in range(3):
        net.apply(r, z, feat, ones, ones, scal, partial, (n, n, n))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        net.apply(r, z, feat, ones, ones, scal, partial, (n, n, n))
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    fl = unet_flops(net, n)
    return {"n": n, "apply_ms": round(ms, 4), "gflop": round(fl / 1e9, 1), "tflops": round(fl / ms / 1e9, 2),
            "mfma_bf16_peak_frac": round(fl / ms / 1e9 / 2500.0, 4)}


def main(argv=None):
    from .nn_precond import CONFIG5_CHANNELS, UNetPrecond, harvest_pairs, train, train_gmg

    p = argparse.ArgumentParser()
    p.add_argument("--out", default=os.path.join(os.path.dirname(os.path.dirname(__file__)), "data",
                                                 "unet_precond.npz"))
    p.add_argument("--n", type=int, default=128)
    p.add_argument("--problems", type=int, default=32)
    p.add_argument("--steps-per-problem", type=int, default=64)
    p.add_argument("--replay", type=int, default=256)
    p.add_argument("--stride", type=int, default=1)
    p.add_argument("--batch", type=int, default=2)
    p.add_argument("--lr", type=float, default=3e-4)
    p.add_argument("--channels", default=",".join(str(c) for c in CONFIG5_CHANNELS))
    p.add_argument("--random-errors", action="store_true",
                   help="train the stand-alone network on random errors (harvest_pairs / train) instead")
    p.add_argument("--per-problem", type=int, default=24, help="random-error samples per problem")
    p.add_argument("--steps", type=int, default=600, help="steps of the random-error training")
    p.add_argument("--eval-n", default="128,256")
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--time-apply", type=int, default=0, help="only time one application at this size with --out weights")
    a = p.parse_args(argv)
    if a.time_apply:
        print(json.dumps(time_apply(a.out, a.time_apply)), flush=True)
        return
    dev = torch.device("cuda:0")
    net = UNetPrecond(tuple(int(c) for c in a.channels.split(",")), device=dev, seed=a.seed)
    print(f"[train_unet] channels {net.channels}, {net.n_params()} parameters", flush=True)
    t0 = time.perf_counter()
    if a.random_errors:
        samples, ops = harvest_pairs(n=a.n, problems=a.problems, per_problem=a.per_problem, device=dev)
        hist = train(net, samples, ops, steps=a.steps, batch=a.batch, lr=a.lr, log=lambda s: print(s, flush=True))
        n_samples = len(samples)
        del samples, ops
        hybrid = False
    else:
        res = train_gmg(net, n=a.n, problems=a.problems, steps_per_problem=a.steps_per_problem, batch=a.batch,
                        lr=a.lr, replay=a.replay, stride=a.stride, seed=a.seed,
                        log=lambda s: print(s, flush=True))
        hist, n_samples = res["loss"], res["samples"]
        hybrid = True
    torch.cuda.synchronize()
    t_t = time.perf_counter() - t0
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    net.save(a.out)
    torch.cuda.empty_cache()
    out = []
    for i, n in enumerate(int(v) for v in a.eval_n.split(",")):
        r = evaluate(net, n, seed=5000 + i, device=dev, hybrid=hybrid)
        print(json.dumps(r), flush=True)
        out.append(r)
    print(json.dumps({"train_s": round(t_t, 1), "samples": n_samples, "final_loss": float(np.mean(hist[-16:])),
                      "params": net.n_params(), "channels": net.channels, "eval": out}), flush=True)


if __name__ == "__main__":
    main()
