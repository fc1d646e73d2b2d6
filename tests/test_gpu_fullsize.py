"""The configurations at their full sizes. The CPU oracle pins the same code paths at
small sizes in test_gpu_lbm.py, test_gpu_physics.py and test_gpu_mg.py; at full size it
checks the link table and two flow steps, and properties that hold at any size check
the rest:

* 256^3 sphere pack (porosity 0.6, MRT, reference Bouzidi, periodic): links, q and the
  Bouzidi weights bit for bit against the oracle's builder (q and the weights on the
  links whose samples stay inside the box), and two fused steps (one of
  each parity) within 1e-5 of the FP32 oracle's reference step; the fused
  in-place step equals the reference's per-kernel sequence (collide -> stream ->
  IBB -> reset, lbm_cuda_ext.py:445-517, run with this library's reference-semantics
  kernels) after an odd and an even number of steps, <= 1e-5 of the field maxima;
  the rest state stays at rest.
* 384^3 gyroid (8 periods of 48 voxels, MRT, Bouzidi): FP16 (BF16) population storage
  stays within 1e-3 (1e-2) of the largest speed from FP32 storage after 100 steps, the
  bounds the 48^3 unit cell meets against the FP64 oracle.
* 192x192x384 pin-fin heat sink (TRT, Zou/He inlet, convective outlet, walls, Bouzidi):
  link table, q and weights bit for bit against the oracle, two steps within 1e-5 of
  the FP32 oracle, and the fused step equal to the per-kernel sequence after 11 and 12
  steps.
* 256^3 blob conduction problem (the bench's): the hierarchy and one V-cycle against
  the oracle's, and the FGMRES(16) solution's FP64 residual with the oracle's operator.
* 512^3 blob conduction problem: GMG-preconditioned flexible GCR(6), flexible CG and
  restarted flexible GMRES(16) reach an FP64 true relative residual of 1e-8 and agree
  on the solution.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DEV = "cuda"


def _feq_t(rho, u, v, w):
    """Second-order D3Q19 equilibrium in FP32 on the GPU (lbm_kernels.cu:77-113)."""
    from fluxcore3d_amd.solver.d3q19 import C, W
    c = torch.as_tensor(C, dtype=torch.float32, device=DEV)
    wq = torch.as_tensor(W, dtype=torch.float32, device=DEV)
    cu = 3.0 * (c[:, 0:1] * u + c[:, 1:2] * v + c[:, 2:3] * w)
    u2 = 1.5 * (u * u + v * v + w * w)
    return wq[:, None] * rho * (1.0 + cu + 0.5 * cu * cu - u2)


def _pack_links(n, seed=0):
    from fluxcore3d_amd import geometry
    phi, solid, _ = geometry.rsa_spheres(n, porosity=0.6, rmin=4.0, rmax=10.0, seed=seed)
    return solid, _links(phi, solid)


def _links(phi, solid):
    from fluxcore3d_amd.pre_post.sdf_tracing import bfl_q_and_weights, build_links_from_mask_gpu
    per = (True, True, True)
    ff, di, _, fff, ok = build_links_from_mask_gpu(torch.as_tensor(solid, device=DEV), periodic=per)
    q, ge, w1, w2 = bfl_q_and_weights(torch.as_tensor(phi, device=DEV), (0.0,) * 3, 1.0, ff, di, periodic=per)
    return dict(fluid_flat=ff, dir_i=di, ff_flat=fff, ff_valid=ok, w1=w1, w2=w2, q_ge_half=ge)


def _perturbed(N, seed, du=0.02):
    g = torch.Generator(device=DEV).manual_seed(seed)
    rho = 1.0 + 0.005 * torch.rand(N, device=DEV, generator=g)
    u, v, w = (du * torch.randn(N, device=DEV, generator=g) for _ in range(3))
    return _feq_t(rho, u, v, w).contiguous()


def _mrt_S(om):
    from tests.cases import mrt_S
    return mrt_S(om)


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))


@pytest.mark.parametrize("steps", [11, 12])
def test_pack_256_fused_step_equals_reference_sequence(steps):
    from fluxcore3d_amd.solver.aa_engine import AAEngine
    from fluxcore3d_amd.solver.lbm_cuda_ext import LBM_CUDA
    n = 256
    N = n ** 3
    solid, L = _pack_links(n)
    om = 1.0 / 0.55
    S = _mrt_S(om).astype(np.float32)
    f0 = _perturbed(N, 1)
    eng = AAEngine(n, n, n, solid, periodic=(1, 1, 1), collision="mrt", omega=om, S=S,
                   links={k: v.cpu().numpy() for k, v in L.items()}, device=DEV)
    eng.load(f0.view(19, n, n, n))
    macro = [torch.empty(N, device=DEV) for _ in range(4)]
    for _ in range(steps):
        eng.step(*macro)
    f_aa = eng.materialize().view(19, N)
    del eng
    # the reference's kernel sequence on f / f_post (reference-semantics kernels of this library)
    f, fp = f0.clone(), torch.empty_like(f0)
    mref = [torch.empty(N, device=DEV) for _ in range(4)]
    st = torch.as_tensor(solid.reshape(-1).astype(np.uint8), device=DEV)
    St = torch.as_tensor(S, device=DEV)
    i32 = lambda t: t.to(torch.int32).contiguous()  # noqa: E731
    from fluxcore3d_amd.solver.d3q19 import OPP
    opp = torch.as_tensor(OPP, device=DEV)
    args = (i32(L["fluid_flat"]), i32(L["dir_i"]), i32(opp[L["dir_i"].long()]), L["w1"].float().contiguous(),
            L["w2"].float().contiguous(), L["q_ge_half"].to(torch.uint8).contiguous(), i32(L["ff_flat"]),
            L["ff_valid"].to(torch.uint8).contiguous())
    for _ in range(steps):
        LBM_CUDA.macro_collide_mrt(f, fp, *mref, st, St)
        LBM_CUDA.stream_pull(fp, f, n, n, n, 1, 1, 1)
        LBM_CUDA.ibb(fp, f, *args)
        LBM_CUDA.reset_solid(f, st)
    assert _rel(macro[0], mref[0]) < 1e-5
    for a in (1, 2, 3):
        assert _rel(macro[a], mref[a]) < 1e-5, a
    fluid = st == 0
    assert _rel(f_aa[:, fluid], f[:, fluid]) < 1e-5


def test_pack_256_links_q_and_two_steps_match_oracle():
    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.pre_post.sdf_tracing import bfl_q_and_weights
    from fluxcore3d_amd.solver.aa_engine import AAEngine
    from oracle import lbm as olbm
    n = 256
    N = n ** 3
    phi, solid, _ = geometry.rsa_spheres(n, porosity=0.6, rmin=4.0, rmax=10.0, seed=0)
    L = _links(phi, solid)
    q, _, _, _ = bfl_q_and_weights(torch.as_tensor(phi, device=DEV), (0.0,) * 3, 1.0, L["fluid_flat"], L["dir_i"],
                                   periodic=(True, True, True))
    links = {k: v.cpu().numpy() for k, v in L.items()}
    ref = olbm.ibb_from_sdf(solid, phi, periodic=(True, True, True))
    assert ref["fluid_flat"].size > 10 ** 6
    for k in ("fluid_flat", "dir_i", "ff_flat", "ff_valid"):
        np.testing.assert_array_equal(links[k], ref[k], err_msg=k)
    # q on every link, the ones that wrap across the periodic faces included: both
    # sides sample on periodic axes at the node's index plus the offset, wrapped
    links["q"] = q.cpu().numpy()
    for k in ("q", "w1", "w2", "q_ge_half"):
        np.testing.assert_array_equal(links[k], ref[k], err_msg=k)
    # the bench's MRT rates and body force, from a perturbed equilibrium
    om = 1.0 / 0.55
    S = np.full(19, 1.2, np.float32)
    S[[0, 3, 5, 7]] = 0.0
    S[[9, 11, 13, 14, 15]] = om
    f0 = _perturbed(N, 3)
    eng = AAEngine(n, n, n, solid, periodic=(1, 1, 1), collision="mrt", omega=om, omega_minus=1.2, S=S,
                   force=(1e-6, 0.0, 0.0), links=links, device=DEV)
    eng.load(f0.view(19, n, n, n))
    macro = [torch.empty(N, device=DEV) for _ in range(4)]
    f = f0.cpu().numpy()
    del f0
    for _ in range(2):
        eng.step(*macro)
        f, *mw = olbm.step(f, solid, n, n, n, (1, 1, 1), "mrt", om, 1.2, S, 0.0, links, force=(1e-6, 0.0, 0.0))
    assert _rel(macro[0].cpu(), torch.from_numpy(mw[0])) < 1e-5
    speed = float(np.sqrt(mw[1] ** 2 + mw[2] ** 2 + mw[3] ** 2).max())
    assert speed > 1e-3
    for a in (1, 2, 3):
        assert float(np.abs(macro[a].cpu().numpy() - mw[a]).max()) / speed < 1e-5, a
    assert _rel(eng.materialize().view(19, N).cpu(), torch.from_numpy(f)) < 1e-5


def test_pack_256_rest_state_stays_at_rest():
    from fluxcore3d_amd.solver.aa_engine import AAEngine
    from fluxcore3d_amd.solver.d3q19 import W
    n = 256
    N = n ** 3
    solid, L = _pack_links(n)
    om = 1.0 / 0.55
    eng = AAEngine(n, n, n, solid, periodic=(1, 1, 1), collision="mrt", omega=om, S=_mrt_S(om),
                   links={k: v.cpu().numpy() for k, v in L.items()}, device=DEV)
    eng.load(torch.as_tensor(np.broadcast_to(W[:, None], (19, N)).copy(), device=DEV, dtype=torch.float32))
    macro = [torch.empty(N, device=DEV) for _ in range(4)]
    for _ in range(40):
        eng.step(*macro)
    assert float((macro[0] - 1.0).abs().max()) < 1e-5
    assert max(float(m.abs().max()) for m in macro[1:]) < 1e-6


def test_gyroid_384_16bit_storage_tracks_fp32():
    """Config 4's 384^3 gyroid: FP16 (BF16) deviation storage within 1e-3 (1e-2) of the
    largest speed from FP32 storage after 100 MRT steps, the bounds the 48^3 unit cell
    meets against the FP64 oracle."""
    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.solver.aa_engine import AAEngine
    n = 384
    N = n ** 3
    phi, solid = geometry.gyroid(n)
    L = {k: v.cpu().numpy() for k, v in _links(phi, solid).items()}
    om = 1.0 / 0.55
    S = _mrt_S(om)
    f0 = _perturbed(N, 2)
    out = {}
    for storage in ("fp32", "fp16", "bf16"):
        eng = AAEngine(n, n, n, solid, periodic=(1, 1, 1), collision="mrt", omega=om, S=S, links=L, storage=storage,
                       device=DEV)
        eng.load(f0.view(19, n, n, n))
        macro = [torch.empty(N, device=DEV) for _ in range(4)]
        for _ in range(100):
            eng.step(*macro)
        out[storage] = torch.stack(macro[1:])
        del eng, macro
        torch.cuda.empty_cache()
    umax = float(out["fp32"].abs().max())
    assert umax > 1e-3
    assert float((out["fp16"] - out["fp32"]).abs().max()) / umax < 1e-3
    assert float((out["bf16"] - out["fp32"]).abs().max()) / umax < 1e-2


def test_blobs_256_hierarchy_vcycle_and_fgmres_against_oracle():
    """The bench's 256^3 blob problem (k_s/k_f = 353): level shapes, masks and inverse
    diagonals, and one V-cycle (face-form coefficients on every level) against the
    oracle's FP32 hierarchy; the GMG-FGMRES(16) solution's FP64 residual measured with
    the oracle's coefficients and operator."""
    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.solver.mg_pcg import fgmres_fp64
    from fluxcore3d_amd.solver.nlevel_mg_cuda import NLevelGeometricMGSolver
    from oracle import mg as omg
    c = geometry.conduction_problem(256, seed=0, device=DEV)
    s = NLevelGeometricMGSolver(torch.device(DEV), torch.float32, c["ax"], c["sign"], c["s"], omega=0.8, n_pre=2,
                                n_post=2, n_coarse=20)
    s.build_if_needed(c["k"], c["rcp"], c["u"], c["v"], c["w"], c["mask"], c["dt"])
    assert all(lv["faces"] is not None for lv in s._levels)
    host = {k: c[k].cpu().numpy() for k in ("k", "rcp", "u", "v", "w", "mask", "b")}
    h = omg.Hierarchy(host["k"], host["rcp"], host["u"], host["v"], host["w"], host["mask"], c["ax"], c["sign"],
                      c["s"], omega=0.8)
    h.set_dt(c["dt"])
    assert [lv["shape"] for lv in s._levels] == [lv["shape"] for lv in h.levels]
    for a, b in zip(s._levels, h.levels):
        np.testing.assert_array_equal(a["mask_f"].cpu().numpy(), b["mask"])
        assert _rel(a["inv_diagM"].cpu(), torch.from_numpy(b["invD"])) < 1e-6
    r = host["b"].astype(np.float32)
    z = s.precondition(torch.as_tensor(r, device=DEV).view(-1)).cpu()
    assert _rel(z.reshape(-1), torch.from_numpy(h.precond(r)).reshape(-1)) < 1e-5
    _, info = fgmres_fp64(s, c["b"], tol=1e-8, max_it=300, m=16)
    assert info["converged"], info["hist"][-3:]
    shape = s._levels[0]["shape"]
    co64 = tuple(a.astype(np.float64) for a in h.levels[0]["co"])
    msk = host["mask"].astype(np.float64)
    bb = host["b"].astype(np.float64) * msk
    x64 = info["x64"].reshape(shape).cpu().numpy()
    res = (bb - omg.mop(x64, co64, c["dt"], c["ax"], h.levels[0]["out_idx"])) * msk
    assert np.linalg.norm(res) / np.linalg.norm(bb) < 1.01e-8


def test_blobs_512_gcr_fcg_and_fgmres_reach_tolerance_and_agree():
    from fluxcore3d_amd import geometry
    from fluxcore3d_amd.solver.mg_pcg import fcg_fp64, fgmres_fp64, gcr_fp64
    from fluxcore3d_amd.solver.nlevel_mg_cuda import NLevelGeometricMGSolver
    c = geometry.conduction_problem(512, seed=1, device=DEV)
    mg = NLevelGeometricMGSolver(torch.device(DEV), torch.float32, c["ax"], c["sign"], c["s"], omega=0.8, n_pre=2,
                                 n_post=2, n_coarse=20)
    mg.build_if_needed(c["k"], c["rcp"], c["u"], c["v"], c["w"], c["mask"], c["dt"])
    xg, gi = gcr_fp64(mg, c["b"], tol=1e-8, max_it=400, m=6)
    assert gi["converged"] and gi["res_rel"] < 1e-8, gi["hist"][-3:]
    hist = np.asarray(gi["hist"][:-1])
    assert np.all(np.diff(hist) <= 1e-12 * hist[:-1])  # minimal residual: never grows
    xg = gi["x64"].clone()
    del gi
    xf, fi = fcg_fp64(mg, c["b"], tol=1e-8, max_it=600)
    assert fi["converged"] and fi["res_rel"] < 1e-8, fi["hist"][-3:]
    m = c["mask"].reshape(-1).double()
    xf64 = fi["x64"].reshape(-1).clone() * m
    del fi
    assert float((xg.reshape(-1) * m - xf64).norm() / xf64.norm()) < 1e-6
    _, ri = fgmres_fp64(mg, c["b"], tol=1e-8, max_it=400, m=16)
    assert ri["converged"] and ri["res_rel"] < 1e-8, ri["hist"][-3:]
    assert float((ri["x64"].reshape(-1) * m - xf64).norm() / xf64.norm()) < 1e-6


def _heat_sink_sim(c, engine):
    from tests.test_gpu_lbm import _sim_for
    return _sim_for(c, "trt", engine, om2=c["omega_minus"])


def test_heat_sink_384_links_q_bit_exact_and_two_steps_match_oracle():
    """Config 3's flow at full size (192x192x384 pin-fin heat sink, TRT magic 1/4, Zou/He
    inlet u = 0.04, convective outlet, walls, Bouzidi links): the GPU link builder and q
    search equal the oracle's bit for bit (every link: the domain is not periodic, both
    clamp), and two fused steps from rest with the planes equal the FP32 oracle's
    reference step to 1e-5."""
    from fluxcore3d_amd.pre_post.sdf_tracing import bfl_q_and_weights, build_links_from_mask_gpu
    from fluxcore3d_amd.solver.lbm_cuda_ext import LBMCUDAUpgrade, flow_bc_plan
    from oracle import lbm as olbm
    from tests import cases
    c = cases.pin_fin_heat_sink()
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    ref = c["links"]
    assert ref["fluid_flat"].size > 10 ** 5
    ff, di, _, fff, ok = build_links_from_mask_gpu(torch.as_tensor(c["solid"], device=DEV))
    q, ge, w1, w2 = bfl_q_and_weights(torch.as_tensor(c["phi"], device=DEV), (0.0,) * 3, 1.0, ff, di)
    for name, got in (("fluid_flat", ff), ("dir_i", di), ("ff_flat", fff), ("ff_valid", ok), ("q", q),
                      ("q_ge_half", ge), ("w1", w1), ("w2", w2)):
        np.testing.assert_array_equal(got.cpu().numpy(), ref[name], err_msg=name)
    sim = _heat_sink_sim(c, "aa")
    LBMCUDAUpgrade.patch(sim)
    plan = flow_bc_plan(sim)
    N = nx * ny * nz
    f = np.broadcast_to(olbm.W32[:, None], (19, N)).copy()
    for _ in range(2):
        sim.step(do_thermal=False)
        f, *mw = olbm.step(f, c["solid_flow"], nx, ny, nz, (0, 0, 0), "trt", c["omega"], c["omega_minus"], None,
                           0.0, ref, None, plan)
    speed = float(np.sqrt(mw[1] ** 2 + mw[2] ** 2 + mw[3] ** 2).max())
    assert speed > 0.03
    for got, want in ((sim.u, mw[1]), (sim.v, mw[2]), (sim.wv, mw[3])):
        assert float(np.abs(got.cpu().numpy().reshape(-1) - want).max()) / speed <= 1e-5
    assert _rel(sim.rho.cpu().reshape(-1), torch.from_numpy(mw[0])) <= 1e-5
    assert _rel(sim.f.cpu().reshape(19, -1), torch.from_numpy(f)) <= 1e-5


@pytest.mark.parametrize("steps", [11, 12])
def test_heat_sink_384_fused_step_equals_reference_sequence(steps):
    """Config 3's flow at full size: the fused in-place step equals the reference's
    per-kernel sequence (collide -> stream -> IBB -> reset -> Zou/He inlet -> convective
    outlet, lbm_cuda_ext.py:445-517) after an odd and an even number of steps, 1e-5 of
    the field maxima."""
    from fluxcore3d_amd.solver.lbm_cuda_ext import LBMCUDAUpgrade
    from tests import cases
    c = cases.pin_fin_heat_sink()
    out = {}
    for engine in ("aa", "reference"):
        sim = _heat_sink_sim(c, engine)
        LBMCUDAUpgrade.patch(sim)
        for _ in range(steps):
            sim.step(do_thermal=False)
        torch.cuda.synchronize()
        out[engine] = [t.detach().clone() for t in (sim.rho, sim.u, sim.v, sim.wv, sim.f)]
        del sim
        torch.cuda.empty_cache()
    for a, b in zip(out["aa"], out["reference"]):
        assert _rel(a, b) <= 1e-5
    assert float(out["aa"][3].abs().max()) > 0.03
