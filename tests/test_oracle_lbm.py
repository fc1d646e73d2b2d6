"""Pins of the LBM oracle (oracle/lbm_oracle.c) — CPU only.

The reference ships no golden vectors (no tests directory); these checks pin
the oracle against its own self-tests (flow_solver_upgrade.py:725-832,
lbm_cuda_ext.py:751-821), closed forms, brute force on tiny grids and library
routines (numpy.linalg.inv, np.roll, torch.linspace).
"""
import numpy as np
import pytest

from oracle import lbm

C = lbm.C_NP


def _rand_f(shape, seed=42, dtype=np.float64):
    """Random positive populations normalised to unit density per node
    (lbm_cuda_ext.py:762-765)."""
    rng = np.random.default_rng(seed)
    f = rng.random((19,) + shape) * 0.1
    return (f / f.sum(0, keepdims=True)).astype(dtype)


def _perturbed_eq(n, seed=7):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.05 * rng.standard_normal(n)
    u = 0.05 * rng.standard_normal((3, n))
    feq = lbm.feq(rho, u[0], u[1], u[2], np.float64)
    return feq * (1.0 + 0.01 * rng.standard_normal(feq.shape))


def test_moment_matrix_is_the_polynomial_basis():
    # rows generated independently from their defining polynomials
    cx, cy, cz = C.T
    c2 = cx**2 + cy**2 + cz**2
    rows = [np.ones(19), 19 * c2 - 30, (21 * c2**2 - 53 * c2 + 24) / 2, cx, (5 * c2 - 9) * cx, cy, (5 * c2 - 9) * cy,
            cz, (5 * c2 - 9) * cz, 3 * cx**2 - c2, (3 * c2 - 5) * (3 * cx**2 - c2), cy**2 - cz**2,
            (3 * c2 - 5) * (cy**2 - cz**2), cx * cy, cy * cz, cx * cz, cx * (cy**2 - cz**2), cy * (cz**2 - cx**2),
            cz * (cx**2 - cy**2)]
    np.testing.assert_array_equal(lbm.moment_matrix(), np.array(rows).astype(int))


def test_minv_identity_and_conserved_rows():
    M = lbm.moment_matrix().astype(np.float64)
    Mi = lbm.minv()
    assert np.abs(M @ Mi - np.eye(19)).max() < 1e-10
    np.testing.assert_array_equal(M[0], 1)
    np.testing.assert_array_equal(M[3], C[:, 0])
    np.testing.assert_array_equal(M[5], C[:, 1])
    np.testing.assert_array_equal(M[7], C[:, 2])


def test_mrt_with_uniform_rates_is_bgk():
    # flow_solver_upgrade.py:766-832: MRT with all non-conserved S = omega is BGK.
    # The reference's FP32 weight literals sum to 1 + 5e-8, so BGK also relaxes
    # the conserved moments by omega*M_c(feq - f); MRT never does. The two
    # differ by exactly M^-1 P_c omega M (f - feq), P_c = rows 0,3,5,7.
    f = _perturbed_eq(20)
    solid = np.zeros(20, np.uint8)
    om = 1.2
    S = np.full(19, om)
    S[[0, 3, 5, 7]] = 0.0
    fb, rho, u, v, w = lbm.macro_collide("bgk", f, solid, om1=om)
    fm, *_ = lbm.macro_collide("mrt", f, solid, S=S)
    M = lbm.moment_matrix().astype(np.float64)
    P = np.zeros((19, 19))
    P[[0, 3, 5, 7], [0, 3, 5, 7]] = 1.0
    dconserved = np.linalg.inv(M) @ P @ (om * (M @ (f - lbm.feq(rho, u, v, w, np.float64))))
    assert np.abs(fm - fb - dconserved).max() < 1e-14
    assert np.abs(dconserved).max() < 1e-8


def test_mrt_conserves_mass_and_momentum():
    f = _perturbed_eq(20, seed=3)
    solid = np.zeros(20, np.uint8)
    S = np.array([0, 1.19, 1.4, 0, 1.4, 0, 1.4, 0, 1.4, 1.8, 1.8, 1.8, 1.8, 1.8, 1.8, 1.8, 1.2, 1.2, 1.2])
    fm, *_ = lbm.macro_collide("mrt", f, solid, S=S)
    assert np.abs(fm.sum(0) - f.sum(0)).max() < 1e-14
    for a in range(3):
        assert np.abs(C[:, a] @ fm - C[:, a] @ f).max() < 1e-14


def test_selftest_mass_conservation_fp32():
    # lbm_cuda_ext.py:751-790 on the FP32 oracle
    f = _rand_f((8 * 10 * 6,), dtype=np.float32)
    solid = np.zeros(f.shape[1], np.uint8)
    fb, rho, *_ = lbm.macro_collide("bgk", f, solid, om1=1.6)
    assert np.abs(fb.sum(0) - f.sum(0)).max() < 1e-5
    S = np.array([0, 1.19, 1.4, 0, 1.4, 0, 1.4, 0, 1.4, 1.8, 1.8, 1.8, 1.8, 1.8, 1.8, 1.8, 1.2, 1.2, 1.2], np.float32)
    fm, *_ = lbm.macro_collide("mrt", f, solid, S=S)
    assert np.abs(fm.sum(0) - f.sum(0)).max() < 1e-5
    assert abs(float(rho.mean()) - 1.0) < 1e-6


def test_trt_equal_rates_is_bgk_and_bgk_fixed_point():
    f = _perturbed_eq(30, seed=5)
    solid = np.zeros(30, np.uint8)
    fb, *_ = lbm.macro_collide("bgk", f, solid, om1=1.3)
    ft, *_ = lbm.macro_collide("trt", f, solid, om1=1.3, om2=1.3)
    assert np.abs(fb - ft).max() < 1e-15
    # omega = 1 lands exactly on the equilibrium of the incoming moments
    f1, rho, u, v, w = lbm.macro_collide("bgk", f, solid, om1=1.0)
    assert np.abs(f1 - lbm.feq(rho, u, v, w, np.float64)).max() < 1e-15


def test_feq_moments_closed_form():
    rng = np.random.default_rng(1)
    rho = 1 + 0.1 * rng.random(50)
    u = 0.05 * rng.standard_normal((3, 50))
    fe = lbm.feq(rho, *u, np.float64)
    wsum = float(lbm.W32.astype(np.float64).sum())  # the reference's literals: 1 + 5e-8
    assert np.abs(fe.sum(0) - rho * wsum).max() < 1e-7
    for a in range(3):
        assert np.abs(C[:, a] @ fe - rho * u[a]).max() < 1e-7


def test_solid_nodes_get_reference_literals():
    f = _rand_f((12,), dtype=np.float32)
    solid = np.zeros(12, np.uint8)
    solid[[1, 5]] = 1
    fp, rho, u, v, w = lbm.macro_collide("bgk", f, solid, om1=1.0)
    np.testing.assert_array_equal(fp[:, 1], lbm.W32)
    assert fp[1, 5] == np.float32(0.05555556) and fp[7, 5] == np.float32(0.02777778)
    assert rho[5] == 1 and u[5] == 0 and v[5] == 0 and w[5] == 0
    r = lbm.reset_solid(f, solid)
    np.testing.assert_array_equal(r[:, 5], lbm.W32)
    np.testing.assert_array_equal(r[:, 0], f[:, 0])


def _stream_brute(f, nx, ny, nz, per):
    out = np.empty_like(f)
    for q in range(19):
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    src = [i - C[q, 0], j - C[q, 1], k - C[q, 2]]
                    own = False
                    for a, n in enumerate((nx, ny, nz)):
                        if per[a]:
                            src[a] %= n
                        elif not 0 <= src[a] < n:
                            own = True
                    s = (i, j, k) if own else tuple(src)
                    out[q, i, j, k] = f[q][s]
    return out


@pytest.mark.parametrize("per", [(1, 1, 1), (0, 0, 0), (0, 1, 1), (1, 0, 1)])
def test_stream_pull_brute_force(per):
    nx, ny, nz = 4, 5, 3
    f = _rand_f((nx, ny, nz), seed=11, dtype=np.float32)
    got = lbm.stream_pull(f, nx, ny, nz, *per).reshape(19, nx, ny, nz)
    np.testing.assert_array_equal(got, _stream_brute(f, nx, ny, nz, per))


def test_stream_periodic_is_roll_and_conserves():
    nx, ny, nz = 6, 5, 4
    f = _rand_f((nx, ny, nz), seed=2)
    got = lbm.stream_pull(f, nx, ny, nz, 1, 1, 1).reshape(19, nx, ny, nz)
    for q in range(19):
        np.testing.assert_array_equal(got[q], np.roll(f[q], tuple(C[q]), axis=(0, 1, 2)))
    assert abs(got.sum() - f.sum()) < 1e-12


def test_ibb_brute_force_and_halfway_case():
    rng = np.random.default_rng(9)
    n = 7 * 6 * 5
    fp = rng.random((19, n)).astype(np.float32)
    fs = rng.random((19, n)).astype(np.float32)
    links = {
        "fluid_flat": np.array([3, 17, 40, 41, 100]), "dir_i": np.array([1, 7, 12, 18, 5]),
        "w1": np.array([1.0, 0.8, 0.6, 0.3, 0.5], np.float32), "w2": np.array([0.0, 0.2, 0.4, 0.7, 0.5], np.float32),
        "q_ge_half": np.array([1, 1, 0, 0, 0], bool), "ff_flat": np.array([0, 0, 12, 60, 7]),
        "ff_valid": np.array([0, 0, 1, 1, 0], bool)}
    got = lbm.ibb(fp, fs, links)
    want = fs.copy()
    for l in range(5):
        x, qi = links["fluid_flat"][l], links["dir_i"][l]
        qo = lbm.OPP[qi]
        a = fp[qo, x]
        if links["q_ge_half"][l]:
            val = links["w1"][l] * a + links["w2"][l] * fp[qi, x]
        elif links["ff_valid"][l]:
            val = links["w1"][l] * a + links["w2"][l] * fp[qo, links["ff_flat"][l]]
        else:
            val = a
        want[qi, x] = val
    np.testing.assert_array_equal(got, want)
    # q = 1/2 (w1 = 1, w2 = 0): halfway bounce-back writes f_post[opp] into dir_i
    assert got[1, 3] == fp[2, 3]


def test_links_brute_force_order():
    rng = np.random.default_rng(4)
    solid = rng.random((5, 4, 6)) < 0.3
    nx, ny, nz = solid.shape
    want = []
    for q in range(1, 19):
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    a, b, c = i + C[q, 0], j + C[q, 1], k + C[q, 2]
                    if solid[i, j, k] or not (0 <= a < nx and 0 <= b < ny and 0 <= c < nz) or not solid[a, b, c]:
                        continue
                    a2, b2, c2 = i - C[q, 0], j - C[q, 1], k - C[q, 2]
                    ok = 0 <= a2 < nx and 0 <= b2 < ny and 0 <= c2 < nz and not solid[a2, b2, c2]
                    want.append(((i * ny + j) * nz + k, q, ok))
    got = lbm.build_links(solid)
    assert list(got["fluid_flat"]) == [w[0] for w in want]
    assert list(got["dir_i"]) == [w[1] for w in want]
    assert list(got["ff_valid"]) == [w[2] for w in want]


def test_linspace_matches_torch():
    torch = pytest.importorskip("torch")
    for steps in (2, 3, 5, 7, 8, 11):
        np.testing.assert_array_equal(lbm.linspace01(steps), torch.linspace(0, 1, steps, dtype=torch.float32).numpy())


def _sphere_phi(n=24, r=6.3, c=(11.4, 12.2, 10.9)):
    i, j, k = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    return (np.sqrt((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2) - r).astype(np.float32)


def test_bfl_q_is_the_root_of_the_trilinear_sdf():
    # Independent root of the trilinear interpolant (scipy map_coordinates,
    # order 1) inside the first coarse +/- bracket of 7 samples
    # (sdf_tracing.py:140-180); agreement within the bisection resolution.
    from scipy.ndimage import map_coordinates

    phi = _sphere_phi()
    links = lbm.ibb_from_sdf(phi < 0, phi)
    x = np.stack(np.unravel_index(links["fluid_flat"], phi.shape), 1).astype(np.float64)
    d = C[links["dir_i"]].astype(np.float64)
    coarse = np.linspace(0.0, 1.0, 7)
    fine = np.linspace(0.0, 1.0, 6 * 4096 + 1)
    pc = map_coordinates(phi.astype(np.float64), (x[:, None, :] + coarse[None, :, None] * d[:, None, :]).T.reshape(3, -1),
                         order=1, mode="nearest").reshape(7, -1).T
    n_checked = 0
    for l in range(len(x)):
        s = pc[l] > 0
        cross = np.nonzero(s[:-1] & ~s[1:])[0]
        if cross.size == 0:
            assert links["q"][l] == np.float32(0.5)
            continue
        a = cross[0]
        tt = fine[(fine >= coarse[a]) & (fine <= coarse[a + 1])]
        pf = map_coordinates(phi.astype(np.float64), (x[l] + tt[:, None] * d[l]).T, order=1, mode="nearest")
        m = np.nonzero(pf <= 0)[0]
        q_ref = tt[m[0]] if m.size else coarse[a + 1]
        assert abs(float(links["q"][l]) - q_ref) < (1.0 / 6) / 1024 + 2.0 / (6 * 4096) + 1e-5, l
        n_checked += 1
    assert n_checked > 0.95 * len(x)


def test_bfl_q_sphere_sanity_and_weights():
    # the interpolated-SDF root scatters around the exact ray/sphere hit
    # (large only on grazing links where |phi| < 0.1 along the link)
    c, r = np.array([11.4, 12.2, 10.9]), 6.3
    phi = _sphere_phi()
    links = lbm.ibb_from_sdf(phi < 0, phi)
    x = np.stack(np.unravel_index(links["fluid_flat"], phi.shape), 1).astype(np.float64)
    d = C[links["dir_i"]].astype(np.float64)
    L = np.linalg.norm(d, axis=1)
    dh = d / L[:, None]
    oc = x - c
    b = np.einsum("ij,ij->i", oc, dh)
    disc = b**2 - (np.einsum("ij,ij->i", oc, oc) - r * r)
    q_exact = (-b - np.sqrt(np.maximum(disc, 0))) / L
    err = np.abs(links["q"] - q_exact)
    assert np.median(err) < 0.02 and np.percentile(err, 90) < 0.1
    assert links["q"].min() > 0 and links["q"].max() <= 1.0
    ge = links["q"] >= 0.5
    np.testing.assert_array_equal(links["q_ge_half"], ge)
    np.testing.assert_array_equal(links["w1"][ge], np.float32(1) / (np.float32(2) * links["q"][ge]))
    np.testing.assert_array_equal(links["w2"][ge], np.float32(1) - links["w1"][ge])
    np.testing.assert_array_equal(links["w1"][~ge], np.float32(2) * links["q"][~ge])
    np.testing.assert_array_equal(links["w2"][~ge], np.float32(1) - np.float32(2) * links["q"][~ge])


def test_zou_he_inlet_and_outlet_moments():
    nx, ny, nz = 6, 5, 4
    rng = np.random.default_rng(3)
    rho = 1 + 0.01 * rng.random((nx, ny, nz))
    u = 0.02 * rng.standard_normal((3, nx, ny, nz))
    f = lbm.feq(rho, *u, np.float64) * (1 + 0.01 * rng.standard_normal((19, nx, ny, nz)))
    solid = np.zeros((nx, ny, nz), np.uint8)
    # velocity inlet at x = 0: the rebuilt node carries exactly u = (u_in, 0, 0)
    g = lbm.zou_he(f, solid, nx, ny, nz, 0, 0, 1, 0, 0.05, 1.0, +1).reshape(19, nx, ny, nz)
    # (the 8-digit literals of W and Q make the regularised part zero-mean
    # only to ~1e-8)
    r = g[:, 0].sum(0)
    assert np.abs((C[:, 0] @ g[:, 0].reshape(19, -1)) / r.reshape(-1) - 0.05).max() < 1e-7
    assert np.abs((C[:, 1] @ g[:, 0].reshape(19, -1))).max() < 1e-7
    np.testing.assert_array_equal(g[:, 1:], f[:, 1:])
    # pressure outlet at x = nx-1: rho = rho_out (or a copy of the interior on backflow)
    h = lbm.zou_he(f, solid, nx, ny, nz, 0, nx - 1, nx - 2, 1, 0.05, 1.0, +1).reshape(19, nx, ny, nz)
    rho_o = h[:, -1].sum(0)
    copied = np.all(h[:, -1] == f[:, -2], axis=0)
    assert np.all(copied | (np.abs(rho_o - 1.0) < 1e-7))


def test_convective_mass_and_copy():
    nx, ny, nz = 5, 4, 3
    f = _rand_f((nx, ny, nz), seed=8) * 1.1
    solid = np.zeros((nx, ny, nz), np.uint8)
    solid[-1, 0, 0] = 1
    g = lbm.convective(f, solid, nx, ny, nz, 0, nx - 1, nx - 2, -1, 1.0).reshape(19, nx, ny, nz)
    fluid = solid[-1] == 0
    assert abs(g[:, -1][:, fluid].sum() / fluid.sum() - 1.0) < 1e-12
    unk = [q for q in range(19) if C[q, 0] == -1]
    ratio = g[unk][:, -1][:, fluid] / f[unk][:, -2][:, fluid]
    assert np.abs(ratio - ratio.flat[0]).max() < 1e-12  # same correction factor everywhere
    np.testing.assert_array_equal(g[:, -1, 0, 0], f[:, -1, 0, 0])  # solid node untouched


# --- Smagorinsky MRT (lbm_kernels.cu:1018-1117) and sponge (:505-528) -------
# Both oracle functions are checked against what the model fixes, computed
# here from first principles: the non-equilibrium stress tensor as an einsum
# over the lattice vectors, its Frobenius norm from numpy.linalg.norm, the
# eddy viscosity as the positive root of its quadratic from numpy.roots, and
# the relaxation rate each moment row actually received, read back from the
# oracle's output through the moment matrix.

S_SMAG = np.array([0, 1.19, 1.4, 0, 1.4, 0, 1.4, 0, 1.4, 1.8, 1.8, 1.8, 1.8, 1.8, 1.8, 1.8, 1.2, 1.2, 1.2])
VISCOUS_ROWS = [9, 11, 13, 14, 15]


def _feq_closed(rho, u):
    """w_i rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u), FP64, with the weight literals."""
    w = lbm.W32.astype(np.float64)[:, None]
    cu = C.astype(np.float64) @ u
    return w * rho * (1 + 3 * cu + 4.5 * cu**2 - 1.5 * (u * u).sum(0))


def _rates_applied(f, f_post, feq):
    """Per-row relaxation rate s_k = [M(f - f_post)]_k / [M(f - feq)]_k."""
    M = lbm.moment_matrix().astype(np.float64)
    return (M @ (f - f_post)) / (M @ (f - feq))


def _smag_state(n=24, seed=21):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.05 * rng.standard_normal(n)
    u = 0.05 * rng.standard_normal((3, n))
    f = _feq_closed(rho, u) * (1.0 + 0.02 * rng.standard_normal((19, n)))
    return f, np.zeros(n, np.uint8)


def test_smagorinsky_with_cs_zero_is_mrt_with_omega0_on_viscous_rows():
    f, solid = _smag_state()
    om0 = 1.7
    fs, *_ = lbm.macro_collide("mrt_smag", f, solid, om1=om0, S=S_SMAG, cs_dx2=0.0)
    S2 = S_SMAG.copy()
    S2[VISCOUS_ROWS] = om0
    fm, *_ = lbm.macro_collide("mrt", f, solid, S=S2)
    assert np.abs(fs - fm).max() < 1e-14
    # the global S of the viscous rows is replaced, not used
    S3 = S_SMAG.copy()
    S3[VISCOUS_ROWS] = 0.3
    fs3, *_ = lbm.macro_collide("mrt_smag", f, solid, om1=om0, S=S3, cs_dx2=0.0)
    assert np.abs(fs3 - fm).max() < 1e-14


@pytest.mark.parametrize("cs_dx2", [1e-3, 1e-2, 0.1])
def test_smagorinsky_rate_is_the_root_of_the_eddy_viscosity_quadratic(cs_dx2):
    f, solid = _smag_state(seed=int(cs_dx2 * 1e4))
    om0 = 1.9
    fs, rho, u, v, w = lbm.macro_collide("mrt_smag", f, solid, om1=om0, S=S_SMAG, cs_dx2=cs_dx2)
    uu = np.stack([u, v, w])
    feq = _feq_closed(rho, uu)
    # rho and u are the node's own moments (u through 1/(rho + 1e-12))
    assert np.abs(rho - f.sum(0)).max() < 1e-14
    assert np.abs(uu - (C.T.astype(np.float64) @ f) / f.sum(0)).max() < 1e-12
    Pi = np.einsum("qn,qa,qb->nab", f - feq, C.astype(np.float64), C.astype(np.float64))
    pif = np.linalg.norm(Pi, "fro", axis=(1, 2))
    nu0 = (1.0 / om0 - 0.5) / 3.0
    cs4 = (1.0 / 3.0) ** 2
    rates = _rates_applied(f, fs, feq)
    clamped = 0
    for n in range(f.shape[1]):
        # nu (nu - nu0) = (Cs dx)^2 |Pi| / (2 rho cs^4)
        roots = np.roots([1.0, -nu0, -cs_dx2 * pif[n] / (2.0 * rho[n] * cs4)])
        nu = float(roots.real.max())
        om = 1.0 / (3.0 * nu + 0.5)
        if om > 1.99:
            om, clamped = float(np.float32(1.99)), clamped + 1
        np.testing.assert_allclose(rates[VISCOUS_ROWS, n], om, rtol=1e-9)
        others = [k for k in range(19) if k not in VISCOUS_ROWS + [0, 3, 5, 7]]
        np.testing.assert_allclose(rates[others, n], S_SMAG[others], rtol=1e-9)
    # the eddy viscosity lowered the rate below omega0 everywhere
    assert rates[9].max() < om0 and clamped == 0


def test_smagorinsky_rate_clamped_at_1_99():
    f, solid = _smag_state(seed=5)
    om0 = 1.995  # nu0 > 0 but 1/(3 nu0 + 1/2) = 1.995 > 1.99
    fs, rho, u, v, w = lbm.macro_collide("mrt_smag", f, solid, om1=om0, S=S_SMAG, cs_dx2=0.0)
    rates = _rates_applied(f, fs, _feq_closed(rho, np.stack([u, v, w])))
    np.testing.assert_allclose(rates[VISCOUS_ROWS], float(np.float32(1.99)), rtol=1e-10)


def _sponge_state(n=40, seed=13):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.05 * rng.standard_normal(n)
    u = 0.05 * rng.standard_normal((3, n))
    f = _feq_closed(rho, u) * (1.0 + 0.05 * rng.standard_normal((19, n)))
    # macros that are not the populations' own: the sponge takes them as given
    rho_s = 1.0 + 0.05 * rng.standard_normal(n)
    u_s = 0.05 * rng.standard_normal((3, n))
    return f, rho_s, u_s


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sponge_sigma_zero_or_negative_is_the_identity(dtype):
    f, rho, u = _sponge_state()
    f = f.astype(dtype)
    sig = np.zeros(f.shape[1])
    sig[::3] = -0.5
    got = lbm.sponge(f, rho, u[0], u[1], u[2], sig)
    np.testing.assert_array_equal(got, f)


def test_sponge_sigma_one_gives_feq_and_partial_sigma_interpolates():
    f, rho, u = _sponge_state()
    feq = _feq_closed(rho, u)
    got = lbm.sponge(f, rho, u[0], u[1], u[2], np.ones(f.shape[1]))
    assert np.abs(got - feq).max() < 1e-15
    sig = np.random.default_rng(2).random(f.shape[1])
    got = lbm.sponge(f, rho, u[0], u[1], u[2], sig)
    assert np.abs(got - ((1 - sig) * f + sig * feq)).max() < 1e-15
    # the sponge pulls the momentum towards rho_s u_s by the factor sigma (to
    # the 5e-8 by which the weight literals miss sum w c c = 1/3)
    j = C.T.astype(np.float64) @ got
    j_want = (1 - sig) * (C.T.astype(np.float64) @ f) + sig * rho * u
    assert np.abs(j - j_want).max() < 1e-7 * np.abs(rho * u).max()


def test_sdf_points_is_the_trilinear_interpolant():
    """sdf_trilinear (sdf_tracing.py:88-125): the trilinear interpolant of phi at grid
    coordinates clamped to [0, n - 1.001] (scipy map_coordinates, order 1, in FP64)."""
    from scipy.ndimage import map_coordinates
    rng = np.random.default_rng(12)
    phi = rng.standard_normal((9, 7, 11)).astype(np.float32)
    origin, pitch = np.array([0.3, -1.2, 2.0], np.float32), 0.5
    g = rng.uniform(-2.0, 12.0, (500, 3))
    x = (origin + pitch * g).astype(np.float32)
    got = lbm.sdf_points(phi, origin, pitch, x)
    gg = (x.astype(np.float64) - origin) / pitch
    gg = np.clip(gg, 0.0, np.array(phi.shape) - 1.001)
    want = map_coordinates(phi.astype(np.float64), gg.T, order=1, mode="nearest")
    assert np.abs(got - want).max() < 1e-5 * np.abs(phi).max()


def test_q_rays_reproduce_the_link_rays_and_weights_closed_form():
    """The generic-ray q search gives the link q bit for bit for the link rays, and
    the weights satisfy w1 = 1/(2q), w2 = 1 - w1 (q >= 1/2), w1 = 2q, w2 = 1 - 2q."""
    phi = _sphere_phi()
    links = lbm.ibb_from_sdf(phi < 0, phi)
    x0 = np.stack(np.unravel_index(links["fluid_flat"], phi.shape), 1).astype(np.float32)
    c = C[links["dir_i"]].astype(np.float32)
    cn = np.sqrt((c * c).sum(1)).astype(np.float32)
    d = (c / (cn + np.float32(1e-30))[:, None]).astype(np.float32)
    q = lbm.q_rays(phi, (0, 0, 0), 1.0, x0, d, cn)
    np.testing.assert_array_equal(q, links["q"])
    qq = np.concatenate([links["q"], np.float32([0.5, 1.0, 0.25, 1e-3])])
    ge, w1, w2 = lbm.bfl_weights(qq)
    np.testing.assert_array_equal(ge, qq >= 0.5)
    assert np.abs(np.where(ge, w1 * 2 * qq, w1 / (2 * qq)) - 1).max() < 1e-6
    assert np.abs(w1 + w2 - 1).max() < 1e-6
    assert w1[-4] == 1 and w2[-4] == 0


@pytest.mark.parametrize("kind", ["bgk", "trt", "mrt"])
def test_guo_force_poiseuille_between_bouzidi_walls(kind):
    """The oracle's Guo body force (a build extension: the reference has none) with
    textbook Bouzidi walls at q = 0.3 and 0.7 (4 x 24 x 4 channel, periodic in x and z,
    FP64, tau = 0.8, F = 1e-5 along x): the steady profile is the parabola
    u(y) = F / (2 nu) (y - y1)(y2 - y) to the discretisation error of the walls
    (5e-3 relative), and the momentum balance F * (fluid nodes) = wall friction holds."""
    nx, ny, nz = 4, 24, 4
    y1, y2 = 0.7, ny - 1 - 0.3
    j = np.arange(ny, dtype=np.float64)
    d = np.minimum(j - y1, y2 - j)
    phi = np.broadcast_to(d[None, :, None], (nx, ny, nz)).astype(np.float32).copy()
    solid = phi < 0
    links = lbm.ibb_from_sdf(solid, phi, periodic=(True, True, True))
    links = dict(links, dir_i=lbm.OPP[links["dir_i"]])  # textbook BFL: the unknown population
    tau = 0.8
    om, nu, F = 1.0 / tau, (tau - 0.5) / 3.0, 1e-5
    om2 = 1.0 / (0.5 + 0.1875 / (tau - 0.5))
    S = np.array([0, 1.19, 1.4, 0, 1.4, 0, 1.4, 0, 1.4, om, 1.4, om, 1.4, om, om, om, 1.98, 1.98, 1.98])
    f = np.broadcast_to(lbm.W32[:, None].astype(np.float64), (19, nx * ny * nz)).copy()
    for _ in range(6000):
        f, rho, u, v, w = lbm.step(f, solid, nx, ny, nz, (1, 1, 1), kind, om, om2, S if kind == "mrt" else None,
                                   0.0, links, force=(F, 0.0, 0.0))
    prof = u.reshape(nx, ny, nz)[1, :, 1]
    fluid = ~solid[1, :, 1]
    exact = F / (2 * nu) * (j - y1) * (y2 - j)
    err = np.linalg.norm(prof[fluid] - exact[fluid]) / np.linalg.norm(exact[fluid])
    assert err < 5e-3, err
    assert prof.max() > 0.5 * exact.max()  # the force drives the flow


@pytest.mark.parametrize("om", [0.7, 1.3, 1.9])
def test_guo_force_mrt_with_uniform_rates_is_bgk_and_adds_f_to_momentum(om):
    """Forced collision (FP64): the post-collision momentum is j + F for BGK, TRT and
    MRT, and MRT with every non-conserved rate equal to omega is forced BGK (the two
    differ only through the weight literals, sum w = 1 + 5e-8)."""
    f = _perturbed_eq(40, seed=11)
    solid = np.zeros(40, np.uint8)
    F = np.array([1e-3, -2e-3, 5e-4])
    S = np.full(19, om)
    S[[0, 3, 5, 7]] = 0.0
    fb, rho, u, v, w = lbm.macro_collide("bgk", f, solid, om1=om, force=F)
    ft, *_ = lbm.macro_collide("trt", f, solid, om1=om, om2=1.1, force=F)
    fm, *_ = lbm.macro_collide("mrt", f, solid, S=S, force=F)
    j = C.T.astype(np.float64) @ f
    lit = 3e-7 * (np.abs(j).max() + np.abs(F).max())  # omega x the 5e-8 by which the weight literals miss
    for g in (fb, ft, fm):
        assert np.abs(C.T.astype(np.float64) @ g - (j + F[:, None])).max() < lit
    # the macros are the Guo velocity (j + F/2) / rho
    assert np.abs(np.stack([u, v, w]) - (j + 0.5 * F[:, None]) / f.sum(0)).max() < 1e-12
    assert np.abs(fm - fb).max() < 1e-7
