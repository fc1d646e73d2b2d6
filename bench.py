"""Benchmark of the fused D3Q19 step on one MI355X.

Workload (flow-only throughput): 256^3 lattice, FP32
populations, MRT (d'Humieres moment basis, s_nu = 1/0.55, all other
non-conserved rates 1.2), Bouzidi links on a random sequential-addition pack
of non-overlapping spheres (radii uniform 4-10, porosity 0.6, seed 0), fully
periodic, Guo body force 1e-6 along x, macros written every step.

Timing: W warm-up steps, then K steps bracketed by synchronisation and HIP
events on the stream the kernels run on. value = fluid-node MLUPS (only
fluid-node updates are counted); roofline.achieved = 152 B x fluid nodes per
launch / average launch time. cpu_baseline: the FP32 oracle (C + OpenMP
restatement of the reference step) on the host cores over a few steps of the
same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_FLOPS = 2.5e15
METRIC = "D3Q19 MLUPS (%HBM roofline) + thermal solve s to 1e-8 resid, NN-PCG vs GMG-PCG"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=50)
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--collision", default="mrt", choices=["bgk", "trt", "mrt", "mrt_smag"])
    p.add_argument("--storage", default="fp32", choices=["fp32", "fp16", "bf16"])
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--layout", default="auto", choices=["auto", "dense", "compact"],
                   help="population layout of the fused step (compact: only fluid and ghost nodes own slots, one "
                        "lane per such node; auto: compact for FP32 storage on this pack, dense for 16-bit)")
    p.add_argument("--cpu-steps", type=int, default=16)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--macro-every-step", type=int, default=1)
    p.add_argument("--also-storage", default="fp16", choices=["fp16", "bf16", "none"],
                   help="also time the same workload with 16-bit population storage (reported beside value)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="run a 256 MiB device-to-device copy first (reference kernel for PMC byte counters)")
    p.add_argument("--thermal-n", type=int, default=256)
    p.add_argument("--no-thermal", action="store_true")
    p.add_argument("--cht-steps", type=int, default=20000,
                   help="coupled heat-transfer steps timed (config 3's length; about 23 s)")
    p.add_argument("--no-cht", action="store_true")
    p.add_argument("--unet-weights", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "fluxcore3d_amd",
                                                          "data", "unet_precond.npz"))
    return p.parse_args()


# Algorithmic HBM bytes per cell of the multigrid and Krylov kernels (FP32 V-cycle,
# FP64 Krylov vectors). Face form: 3 face-coefficient arrays (12 B) instead of the 6
# coefficients + diagL (28 B) per cell.
#   V-cycle level (not the coarsest): first two sweeps from x = 0 (b, invD, coefficients,
#   write x), residual + restriction (x, b, mask, coefficients, 1/8 coarse write),
#   prolongation + correction (1/8 coarse read, x, mask, write x), two sweeps (x,
#   coefficients, b, invD, write): 24 + 24.5 + 12.5 + 2 * 28 = 117 B (face form),
#   40 + 40.5 + 12.5 + 2 * 44 = 181 B (per-cell arrays); coarsest level: one pass of
#   b, invD, coefficients and x (24 / 40 B), its 20 sweeps run in LDS.
#   Krylov work per iteration on the finest level: FCG dot(z, Ap_old) 12 + new direction
#   20 + A p with p.Ap, p.r 40 (56) + x, r update 52 = 124 (140) B; GCR(6) 280 (296) B;
#   FGMRES(16) 184 (200) B averaged over a restart cycle. True residual: 44 (60) B, twice
#   per solve.
_VC_LEVEL_B = (117.0, 181.0)
_VC_COARSE_B = (24.0, 40.0)
_KRYLOV_B = {"fcg": (124.0, 140.0), "gcr6": (280.0, 296.0), "fgmres16": (184.0, 200.0)}
_TRUE_RES_B = (44.0, 60.0)


def _form(lv):
    return 0 if lv.get("faces") is not None else 1


def vcycle_bytes(levels) -> float:
    """Algorithmic bytes of one V-cycle preconditioner application."""
    total = 0.0
    for i, lv in enumerate(levels):
        n = float(np.prod(lv["shape"]))
        total += n * (_VC_COARSE_B if i == len(levels) - 1 else _VC_LEVEL_B)[_form(lv)]
    return total


# one A z in FP32 (z read, face coefficients or the seven per-cell values, Lo written)
# plus the scaled residual (r, Lo, mask, d read, one value written): the extra fine-level
# work of a hybrid direction besides its V-cycle and the network
_HYBRID_B = (4.0 + 12 + 4 + 20, 4.0 + 28 + 4 + 20)  # face form, per-cell arrays


def solve_roofline(levels, iterations, solver, measured_s, network_flops=0.0, network_iterations=0,
                   hybrid=False):
    """T_roof = (iterations x (V-cycle + Krylov bytes) + 2 true residuals) / 8 TB/s, plus
    the U-Net's FLOPs / 2.5 PFLOP/s for iterations that apply the network (instead of
    the V-cycle, or with ``hybrid`` after it, with the residual the V-cycle leaves);
    frac = T_roof / T_measured."""
    n0 = float(np.prod(levels[0]["shape"]))
    f0 = _form(levels[0])
    vc = vcycle_bytes(levels)
    kb = _KRYLOV_B[solver][f0] * n0
    mg_it = iterations - network_iterations
    nn_b = (vc + _HYBRID_B[f0] * n0) if hybrid else 0.0
    bytes_ = mg_it * (vc + kb) + network_iterations * (kb + nn_b) + 2 * _TRUE_RES_B[f0] * n0
    t_roof = bytes_ / (HBM_PEAK_GBS * 1e9) + network_iterations * network_flops / MFMA_BF16_PEAK_FLOPS
    return {"t_roof_s": round(t_roof, 5), "t_measured_s": round(measured_s, 5),
            "frac": round(t_roof / max(measured_s, 1e-12), 4), "algorithmic_bytes": int(bytes_),
            "vcycle_bytes": int(vc), "krylov_bytes_per_iteration": int(kb)}


def thermal_cpu_baseline(n=64, seed=0):
    """The oracle's FP64-outer flexible CG with the FP32 V-cycle (NumPy, one core) on the
    n^3 version of the blob problem, to a relative residual of 1e-8."""
    from fluxcore3d_amd import geometry
    from oracle import mg as omg

    c = geometry.conduction_problem(n, seed=seed, device="cpu")
    f32 = lambda t: t.cpu().numpy().astype(np.float32)  # noqa: E731
    h = omg.Hierarchy(f32(c["k"]), f32(c["rcp"]), f32(c["u"]), f32(c["v"]), f32(c["w"]), f32(c["mask"]) > 0,
                      c["ax"], c["sign"], c["s"], omega=0.8)
    h.set_dt(c["dt"])
    from threadpoolctl import threadpool_limits

    with threadpool_limits(1):  # one core: NumPy's BLAS calls (the dot products) included
        t0 = time.perf_counter()
        _, info = omg.fcg_mg_fp64(h, f32(c["b"]), tol=1e-8, max_it=500)
        secs = time.perf_counter() - t0
    return {"n": n, "s": secs, "iterations": info["it"], "res": info["hist"][-1]}


def thermal_leg(n, device, seed=0, weights=None):
    """GMG-preconditioned flexible CG (FP32 V-cycle, FP64 outer loop, the reference's
    PCG), GMG-preconditioned flexible GCR(6), restarted flexible GMRES(16) and the
    reference's BiCGSTAB, and flexible GCR(6) with the U-Net preconditioner (V-cycle
    fallback), all to a relative residual of 1e-8 on the same 256^3 blob conduction
    problem."""
    import torch

    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.solver.mg_pcg import FCGWorkspace, fcg_fp64
    from fluxcore3d_amd.solver.nlevel_mg_cuda import NLevelGeometricMGSolver

    c = geometry.conduction_problem(n, seed=seed, device=device)
    mg = NLevelGeometricMGSolver(device, torch.float32, c["ax"], c["sign"], c["s"], omega=0.8, n_pre=2, n_post=2,
                                 n_coarse=20)
    mg.build_if_needed(c["k"], c["rcp"], c["u"], c["v"], c["w"], c["mask"], c["dt"])
    ws = FCGWorkspace(c["k"].shape, device)
    fcg_fp64(mg, c["b"], tol=1e-8, max_it=3, ws=ws)  # warm-up (allocations, first launches)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, info = fcg_fp64(mg, c["b"], tol=1e-8, max_it=500, ws=ws)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    from fluxcore3d_amd.solver.mg_pcg import GCRWorkspace, gcr_fp64

    gws = GCRWorkspace(c["k"].shape, device, 6)
    gcr_fp64(mg, c["b"], tol=1e-8, max_it=3, m=6, ws=gws)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xg, gi = gcr_fp64(mg, c["b"], tol=1e-8, max_it=500, m=6, ws=gws)
    torch.cuda.synchronize()
    gcr_secs = time.perf_counter() - t0
    del gws
    from fluxcore3d_amd.solver.mg_pcg import FGMRESWorkspace, fgmres_fp64

    fws = FGMRESWorkspace(c["k"].shape, device, 16)
    # warm-up: one full cycle launches the Arnoldi kernels of every basis count once
    fgmres_fp64(mg, c["b"], tol=1e-8, max_it=16, m=16, ws=fws)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xf, fi = fgmres_fp64(mg, c["b"], tol=1e-8, max_it=500, m=16, ws=fws)
    torch.cuda.synchronize()
    fgmres_secs = time.perf_counter() - t0
    del fws
    from fluxcore3d_amd.solver.mg_pcg import BiCGSTABWorkspace, bicgstab_mg

    bws = BiCGSTABWorkspace(c["k"].shape, device)
    bicgstab_mg(mg, c["b"], tol=1e-8, max_it=2, ws=bws)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, bi = bicgstab_mg(mg, c["b"], tol=1e-8, max_it=500, ws=bws)
    torch.cuda.synchronize()
    bicg_secs = time.perf_counter() - t0
    del bws
    out = {"problem": f"{n}^3 smooth-noise blobs (seed {seed}), k_s/k_f = {c['ratio']:.3g}, dt = {c['dt']:g}",
           "levels": [list(lv["shape"]) for lv in mg._levels], "gmg_pcg_s": round(secs, 4),
           "gmg_pcg_iterations": info["it"], "final_rel_residual_fp64": info["res_rel"],
           "converged": bool(info["converged"]), "gmg_gcr6_s": round(gcr_secs, 4),
           "gmg_gcr6_iterations": gi["it"], "gmg_gcr6_final_rel_residual_fp64": gi["res_rel"],
           "gmg_fgmres16_s": round(fgmres_secs, 4), "gmg_fgmres16_iterations": fi["it"],
           "gmg_fgmres16_final_rel_residual_fp64": fi["res_rel"], "gmg_fgmres16_converged": bool(fi["converged"]),
           "gmg_fgmres16_vs_gmg_pcg_speedup": round(secs / max(fgmres_secs, 1e-9), 3),
           # the reference's patch_mg_pcg picks BiCGSTAB here (0.1 max(diagL) above 1.5); its tolerance is on the
           # recursive residual, as in the reference
           "gmg_bicgstab_s": round(bicg_secs, 4), "gmg_bicgstab_iterations": bi["it"],
           "gmg_bicgstab_converged": bool(bi["converged"]), "gmg_bicgstab_final_rel_residual": bi["res_rel"],
           "nn_pcg_s": None}
    lv = mg._levels
    out["roofline"] = {
        "bound": "hbm", "peak_gbs": HBM_PEAK_GBS,
        "gmg_pcg": solve_roofline(lv, info["it"], "fcg", secs),
        "gmg_gcr6": solve_roofline(lv, gi["it"], "gcr6", gcr_secs),
        "gmg_fgmres16": solve_roofline(lv, fi["it"], "fgmres16", fgmres_secs),
        "model": "T_roof = sum over launches of algorithmic bytes / 8 TB/s (V-cycle 117 B per cell per level in "
                 "face form, Krylov bytes per iteration as in bench.py); frac = T_roof / T_measured"}
    out["cht"] = cht_leg(device)
    if weights is None or not os.path.exists(weights):
        out["note"] = "no trained U-Net weights found; NN-PCG not run"
        return out
    from fluxcore3d_amd.solver.nn_precond import UNetPrecond, nn_fcg

    net = UNetPrecond.load(weights, device=device)
    nn_fcg(mg, net, c["b"], c["k"], tol=1e-8, max_it=2)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xn, ni = nn_fcg(mg, net, c["b"], c["k"], tol=1e-8, max_it=500)
    torch.cuda.synchronize()
    nn_secs = time.perf_counter() - t0
    flops = float(net.flops(tuple(c["k"].shape))) if hasattr(net, "flops") else 0.0
    out["roofline"]["nn_gcr"] = solve_roofline(lv, ni["it"], "gcr6", nn_secs, network_flops=flops,
                                               network_iterations=int(ni.get("it_nn") or 0), hybrid=True)
    out.update({"nn_pcg_s": round(nn_secs, 4), "nn_pcg_iterations": ni["it"],
                "nn_iterations_with_network": ni.get("it_nn"), "nn_fallback_to_gmg": bool(ni["fallback"]),
                "nn_final_rel_residual_fp64": ni["res_rel"], "nn_converged": bool(ni["converged"]),
                "nn_weights": os.path.relpath(weights, os.path.dirname(os.path.abspath(__file__))),
                "nn_vs_gmg_speedup": round(secs / max(nn_secs, 1e-9), 3),
                "nn_mode": "hybrid: each flexible-GCR(6) direction is one V-cycle plus the network's correction "
                           "of the residual the V-cycle leaves",
                "nn_network": {"channels": net.channels, "parameters": net.n_params(),
                               "gflop_per_application": round(flops / 1e9, 1)},
                "nn_vs_gmg_gcr6_iterations": round(ni["it"] / max(gi["it"], 1), 3),
                # the same Krylov method with the V-cycle alone is the fair yardstick of the
                # network's part: GCR(6) by itself already halves GMG-PCG's iterations here
                "nn_speedup_source": ("GMG-GCR(6) fallback after the network stalled" if ni["fallback"] else
                                      "network" if ni["it"] < gi["it"] else
                                      "flexible GCR(6) (the network does not bring the iterations below "
                                      "GMG-GCR(6)'s)"),
                "gmg_gcr6_vs_gmg_pcg_speedup": round(secs / max(gcr_secs, 1e-9), 3)})
    return out


def cht_leg(device):
    """The coupled heat-sink conduction-advection system at full size (pin-fin plate,
    192x192x384, upwind advection, varying rho*c_p: non-symmetric): flexible CG (the
    reference's PCG, reported whether or not it converges), the reference's BiCGSTAB
    branch, GMG-GCR(6) and GMG-FGMRES(16) to an FP64 relative residual of 1e-8."""
    import math

    import torch

    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.solver.mg_pcg import bicgstab_mg, fcg_fp64, fgmres_fp64, gcr_fp64
    from fluxcore3d_amd.solver.nlevel_mg_cuda import NLevelGeometricMGSolver

    c = geometry.cht_problem(scale=1.0, device=device)
    mg = NLevelGeometricMGSolver(device, torch.float32, c["ax"], c["sign"], c["s"], omega=0.8, n_pre=2, n_post=2,
                                 n_coarse=20)
    mg.build_if_needed(c["k"], c["rcp"], c["u"], c["v"], c["w"], c["mask"], c["dt"])
    out = {"problem": "pin-fin heat sink 192x192x384, k_s/k_f = 1000, rho*c_p ratio 2000, channel flow 0.04 along z "
                      "(synthetic), dt = 1000", "face_form": mg._levels[0]["faces"] is not None}
    for name, fn, kw, cap in (("gmg_pcg", fcg_fp64, {}, 100), ("gmg_bicgstab", bicgstab_mg, {}, 400),
                              ("gmg_gcr6", gcr_fp64, {"m": 6}, 400), ("gmg_fgmres16", fgmres_fp64, {"m": 16}, 400)):
        fn(mg, c["b"], tol=1e-8, max_it=max(2, kw.get("m", 0)), **kw)  # warm-up: a full cycle of every kernel
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, info = fn(mg, c["b"], tol=1e-8, max_it=cap, **kw)
        torch.cuda.synchronize()
        res = info["res_rel"]
        out.update({f"{name}_s": round(time.perf_counter() - t0, 4), f"{name}_iterations": info["it"],
                    f"{name}_converged": bool(info["converged"]),
                    f"{name}_final_rel_residual_fp64": res if math.isfinite(res) else None})
        del info
    return out


D3Q7_FLUID_BYTES = 7 * 4 * 2 + 3 * 4 + 2 + 4  # populations in and out, velocity, flag and neighbour mask, T out
D3Q7_SOLID_BYTES = 2 + 4 + 4  # flag and neighbour mask, solid temperature in, temperature out
D3Q7_INTERFACE_BYTES = 8  # energy accumulator read and written (solid nodes next to fluid)


def coupled_leg(device, steps, warmup, every=50):
    """Config 3's coupled conjugate heat transfer at full size: 192x192x384 pin-fin
    heat sink, TRT (magic 1/4) flow from rest with Zou/He inlet u = 0.04 and convective
    outlet, D3Q7 fluid temperature every step, implicit solid conduction (k_s/k_f =
    1000, rho_cp ratio 1.6, base flux as 1e-3 K per step on the plate's bottom layer)
    by GMG-preconditioned FCG to 1e-8 every ``every`` steps. Reports the whole loop's
    time per step and, timed on their own, the flow step, the D3Q7 kernel (with its
    HBM roofline) and the solid solves."""
    import types

    import torch

    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.pre_post.sdf_tracing import build_ibb_from_sdf_gpu
    from fluxcore3d_amd.solver.cht_d3q7 import FLUID, SOLID, WALL, CoupledCHT
    from fluxcore3d_amd.solver.flow_state import FlowState
    from fluxcore3d_amd.solver.lbm_cuda_ext import LBMCUDAUpgrade

    phi, solid, solid_flow = geometry.pin_fin_plate()
    solid_flow[:, 0] = True
    nx, ny, nz = solid.shape
    tau = 0.6
    cfg = types.SimpleNamespace(flow_dir="+Z", domain_walls=["+X", "-X", "+Y", "-Y"], domain_outlets=[],
                                transverse_walls=False, outlet_bc_mode="convective", rho_out_lat=1.0, u_lat_max=0.12,
                                collision="trt", sponge_length=0, u_in_lat=0.04, fc_engine="aa")
    sim = FlowState(nx, ny, nz, cfg, solid, solid_flow, phi, 1.0 / tau, device,
                    omega_minus=1.0 / (0.5 + 0.25 / (tau - 0.5)))
    build_ibb_from_sdf_gpu(sim, (0.0, 0.0, 0.0), 1.0, (0, 0, 0))
    LBMCUDAUpgrade.patch(sim)
    src = torch.zeros((nx, ny, nz), device=device)
    src[:, 0, :] = 1e-3
    src[~sim.solid] = 0.0
    cht = CoupledCHT(sim, src=src, k_ratio=1000.0, rho_cp_ratio=1.6, couple_every=every)
    for _ in range(warmup):
        cht.step()
    cht.sync()  # the thermal stream's work joins the current stream before each event
    torch.cuda.synchronize()
    solves0 = len(cht.history)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(steps):
        cht.step()
    cht.sync()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1)
    hist = cht.history[solves0:]
    fl = cht.flag.reshape(-1)
    kind = fl & 3
    # the state after the coupled run (the timings below step the parts on their own)
    T = cht.temperature()
    temps = {"T_fluid_max": round(float(T.reshape(-1)[kind == FLUID].max()), 6),
             "T_solid_max": round(float(cht.T_solid.max()), 6),
             "T_outlet_mean": round(float(T[:, :, -1][(kind.view(nx, ny, nz)[:, :, -1] == FLUID)].mean()), 6)}
    n_fluid = int((kind == FLUID).sum())
    n_solid = int((kind == SOLID).sum())
    n_wall = int((kind == WALL).sum())
    # interface solid nodes (a fluid face neighbour): from the accumulator pattern of one step
    cht.acc.zero_()
    cht.transport.step(sim.u, sim.v, sim.wv, cht.T_solid, acc=cht.acc)
    n_iface = int((cht.acc != 0).sum())
    # flow step, D3Q7 kernel and solid solve on their own (events on torch's current stream,
    # the stream every launch here uses)
    k = 200
    ev0.record()
    for _ in range(k):
        sim.step(do_thermal=False)
    ev1.record()
    torch.cuda.synchronize()
    flow_ms = ev0.elapsed_time(ev1) / k
    ev0.record()
    for _ in range(k):
        cht.transport.step(sim.u, sim.v, sim.wv, cht.T_solid, acc=cht.acc)
    ev1.record()
    torch.cuda.synchronize()
    d3q7_ms = ev0.elapsed_time(ev1) / k
    d3q7_bytes = (n_fluid * D3Q7_FLUID_BYTES + n_solid * D3Q7_SOLID_BYTES + n_iface * D3Q7_INTERFACE_BYTES
                  + n_wall * 2)
    ev0.record()
    info = cht.solid.step(cht.T_solid, cht.acc, cht.src)
    ev1.record()
    torch.cuda.synchronize()
    solve_ms = ev0.elapsed_time(ev1)
    return {
        "workload": f"pin-fin heat sink {nx}x{ny}x{nz}: TRT flow + D3Q7 temperature every step, solid conduction "
                    f"(k_s/k_f 1000, rho_cp 1.6) by GMG-FCG to 1e-8 every {every} steps; base flux 1e-3 K/step",
        "steps": steps, "warmup": warmup, "ms_per_step": round(ms / steps, 5),
        "thermal_stream_overlap": cht.overlap,
        "mlups_fluid": round(n_fluid * steps / (ms * 1e-3) / 1e6, 2), "wall_s_timed": round(wall, 3),
        "fluid_nodes": n_fluid, "solid_nodes": n_solid, "interface_solid_nodes": n_iface,
        "solid_solves": len(hist), "solid_solve_iterations": [h["it"] for h in hist][:10],
        "solid_solve_it_mean": round(float(np.mean([h["it"] for h in hist])), 2) if hist else None,
        "solid_solve_max_res": max(h["res_rel"] for h in hist) if hist else None,
        "flow_ms_per_step": round(flow_ms, 5),
        "d3q7": {"ms": round(d3q7_ms, 5), "algorithmic_bytes": d3q7_bytes,
                 "achieved_gbs": round(d3q7_bytes / (d3q7_ms * 1e-3) / 1e9, 1),
                 "roofline_frac": round(d3q7_bytes / (d3q7_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "solid_solve_ms": round(solve_ms, 3), "solid_solve_it": info["it"],
        "box": list(cht.solid.box), "box_unknowns": cht.solid.n_unknowns, **temps,
    }


def half_storage_leg(args, solid, links, S, omega, macro, dev, n_fluid):
    """The same workload with 16-bit populations (dense storage): a secondary line, timed
    over at least 200 steps after at least 30 warm-up steps."""
    import torch
    N = args.n ** 3
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    from fluxcore3d_amd.solver.aa_engine import AAEngine
    from fluxcore3d_amd.solver.d3q19 import W

    eng16 = AAEngine(args.n, args.n, args.n, solid, periodic=(True, True, True), collision=args.collision,
                     omega=omega, omega_minus=1.2, S=S, force=(1e-6, 0.0, 0.0), links=links,
                     storage=args.also_storage, device=dev)
    eng16.load(torch.as_tensor(np.broadcast_to(W[:, None], (19, N)).copy(), device=dev))
    # past the step-time transient after loading (the first ~20 steps run up to 10 % slower)
    w16 = max(args.warmup, 30)
    for _ in range(w16):
        eng16.step(*macro)
    torch.cuda.synchronize()
    k16 = max(args.steps // 5, 200)
    ev0.record()
    for _ in range(k16):
        eng16.step(*macro)
    ev1.record()
    torch.cuda.synchronize()
    ms16 = ev0.elapsed_time(ev1) / k16
    half = {"storage": args.also_storage, "steps": k16, "warmup": w16, "ms_per_step": round(ms16, 5),
            "mlups": round(n_fluid / (ms16 * 1e-3) / 1e6, 2),
            "roofline_frac": round(eng16.bytes_per_update() * n_fluid / (ms16 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_update": eng16.bytes_per_update(), "lanes": eng16.lanes}
    del eng16
    torch.cuda.empty_cache()
    return half


def build_workload(n, seed, collision, storage, device, layout="auto"):
    import torch

    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.pre_post.sdf_tracing import bfl_q_and_weights, build_links_from_mask_gpu
    from fluxcore3d_amd.solver.aa_engine import AAEngine
    from fluxcore3d_amd.solver.d3q19 import W

    t0 = time.perf_counter()
    phi, solid, info = geometry.rsa_spheres(n, porosity=0.6, rmin=4.0, rmax=10.0, seed=seed)
    per = (True, True, True)
    solid_t = torch.as_tensor(solid, device=device)
    ff, di, do, fff, ok = build_links_from_mask_gpu(solid_t, periodic=per)
    q, ge, w1, w2 = bfl_q_and_weights(torch.as_tensor(phi, device=device), (0.0, 0.0, 0.0), 1.0, ff, di,
                                      periodic=per)
    links = {k: v.cpu().numpy() for k, v in dict(fluid_flat=ff, dir_i=di, ff_flat=fff, ff_valid=ok, w1=w1, w2=w2,
                                                   q_ge_half=ge).items()}
    omega = 1.0 / 0.55
    S = np.full(19, 1.2, np.float32)
    S[[0, 3, 5, 7]] = 0.0
    S[[9, 11, 13, 14, 15]] = omega
    eng = AAEngine(n, n, n, solid, periodic=per, collision=collision, omega=omega, omega_minus=1.2, S=S,
                   force=(1e-6, 0.0, 0.0), links=links, storage=storage, device=device, layout=layout)
    f0 = torch.as_tensor(np.broadcast_to(W[:, None], (19, n ** 3)).copy(), device=device)
    eng.load(f0)
    setup_s = time.perf_counter() - t0
    return eng, solid, links, S, omega, info, setup_s


def cpu_baseline(solid, links, S, omega, collision, steps):
    """The FP32 oracle step (C + OpenMP restatement of the reference step, with the same
    Guo body force) on the host, on all cores and on one core (kind "port")."""
    from oracle import lbm as olbm

    n = solid.shape[0]
    N = solid.size
    n_fluid = int((~solid).sum())
    threads = int(olbm._lib(np.float32).orc_num_threads())
    out = {}
    for label, nthr, k in (("all", threads, steps), ("one", 1, max(2, steps // 8))):
        olbm.set_threads(nthr)
        f = np.broadcast_to(olbm.W32[:, None], (19, N)).copy()
        if label == "all":  # warm-up (the 1-thread sample follows on warm caches)
            f, *_ = olbm.step(f, solid, n, n, n, (1, 1, 1), collision, omega, 1.2, S, 0.0, links,
                              force=(1e-6, 0.0, 0.0))
        t0 = time.perf_counter()
        for _ in range(k):
            f, *_ = olbm.step(f, solid, n, n, n, (1, 1, 1), collision, omega, 1.2, S, 0.0, links,
                              force=(1e-6, 0.0, 0.0))
        dt = time.perf_counter() - t0
        out[label] = (n_fluid * k / dt / 1e6, k, dt)
    olbm.set_threads(threads)
    (v_all, k_all, t_all), (v_one, k_one, t_one) = out["all"], out["one"]
    return {"value": round(v_all, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "value_1_thread": round(v_one, 3),
            "sample": f"{k_all} steps (+1 warm-up) of the same {n}^3 {collision} workload with the body force, FP32 "
                      f"oracle (C/OpenMP restatement of the reference step) on {threads} threads, {t_all:.1f} s; "
                      f"{k_one} steps on 1 thread (no warm-up step), {t_one:.1f} s"}


def traffic_per_launch(args, layout):
    """HBM bytes per fused-step launch (FETCH_SIZE + WRITE_SIZE, each scaled by its
    reading for a copy of known size) from the committed rocprofv3 PMC summary of
    this workload, or None when no summary matches."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "aa_step_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        table = json.load(fh)
    key = f"{args.n}^3 {args.collision} {args.storage}" + (" compact" if layout == "compact" else "")
    rec = table.get(key)
    return None if rec is None else int(rec["fetch_bytes"] + rec["write_bytes"])


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU: run it with --gpus 1 (the path is not sharded "
                                   "across GPUs)", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("bench.py needs a HIP device")
    dev = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones(64 * 1024 * 1024, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    eng, solid, links, S, omega, info, setup_s = build_workload(args.n, args.seed, args.collision, args.storage, dev,
                                                                args.layout)
    N = args.n ** 3
    n_fluid = int((~solid).sum())
    macro = [torch.empty(N, device=dev) for _ in range(4)]
    # The 16-bit secondary line runs first: after the host-side set-up the GPU's clocks take
    # tens of milliseconds of load to settle, and the headline's warm-up steps alone leave
    # about 2 % of that ramp inside a 20-step timed region.
    half = None
    if args.also_storage != "none" and args.storage == "fp32":
        half = half_storage_leg(args, solid, links, S, omega, macro, dev, n_fluid)
    for _ in range(args.warmup):
        eng.step(*macro)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    ev0.record()
    for s in range(args.steps):
        if args.macro_every_step and s % args.macro_every_step == 0:
            eng.step(*macro)
        else:
            eng.step()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = ev0.elapsed_time(ev1)
    ms_step = ms / args.steps
    mlups = n_fluid * args.steps / (ms * 1e-3) / 1e6
    bpu = eng.bytes_per_update()
    achieved = bpu * n_fluid / (ms_step * 1e-3) / 1e9
    ok = bool(torch.isfinite(macro[1]).all()) and bool(torch.isfinite(macro[0]).all())
    ux_mean = float(macro[1][torch.as_tensor(~solid.reshape(-1), device=dev)].mean())
    n_links = int(eng.n_links)
    layout, lanes = eng.layout, eng.lanes
    thermal = None
    if not args.no_thermal:
        thermal = thermal_leg(args.thermal_n, dev, weights=args.unet_weights)
    coupled = None
    if not args.no_cht:
        torch.cuda.empty_cache()
        coupled = coupled_leg(dev, args.cht_steps, warmup=100)
        torch.cuda.empty_cache()
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(solid, links, S, omega, args.collision, args.cpu_steps)
        if thermal is not None:
            tc = thermal_cpu_baseline(64)
            it_gpu = thermal["gmg_pcg_iterations"]
            scale = (args.thermal_n / tc["n"]) ** 3 * it_gpu / tc["iterations"]
            thermal["cpu_baseline"] = {
                "value": round(tc["s"] * scale, 2), "unit": "s", "cores": 1, "kind": "port",
                "extrapolated": True,
                "sample": f"oracle FP64-outer flexible CG with the FP32 V-cycle (NumPy, one core) on the "
                          f"{tc['n']}^3 blob problem: {tc['s']:.2f} s, {tc['iterations']} iterations to "
                          f"{tc['res']:.2e}; value = that time x ({args.thermal_n}/{tc['n']})^3 cells x "
                          f"{it_gpu}/{tc['iterations']} iterations (extrapolated to the {args.thermal_n}^3 solve)",
                "measured_64cube_s": round(tc["s"], 3)}
    line = {
        "metric": METRIC,
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.storage,
        "data": "synthetic",
        "config": {"workload": f"flow-only throughput: {args.n}^3 {args.collision.upper()} + Bouzidi, periodic RSA "
                               f"sphere pack porosity {info['porosity']:.4f} (seed {args.seed}), body force 1e-6 x",
                   "grid": [args.n] * 3, "fluid_nodes": n_fluid, "n_links": n_links,
                   "storage": args.storage, "collision": args.collision, "ibb": "reference Bouzidi (periodic links)",
                   "layout": layout, "lanes": lanes,
                   "macro_every_step": args.macro_every_step},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_per_launch(args, layout),
                     "algorithmic_bytes_per_update": bpu,
                     "algorithmic_bytes_per_launch": bpu * n_fluid},
        "cpu_baseline": cpu,
        "half_storage": half,
        "thermal": thermal,
        "coupled_cht": coupled,
        "mlups_all_nodes": round(N * args.steps / (ms * 1e-3) / 1e6, 2),
        "wall_s_timed": round(wall, 4),
        "setup_s": round(setup_s, 2),
        "finite": ok,
        "ux_mean": ux_mean,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
