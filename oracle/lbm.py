"""ctypes face of oracle/lbm_oracle.c plus the integer Bouzidi link builder.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py). Each function names the
reference code it restates (paths relative to the fluxcore3D tree).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1])
C_NP = np.stack([CX, CY, CZ], 1)
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15])
W32 = np.array([0.33333333] + [0.05555556] * 6 + [0.02777778] * 12, dtype=np.float32)

_libs = {}


def build() -> None:
    """Compile the oracle libraries with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _lib(dtype):
    key = np.dtype(dtype).name
    if key not in _libs:
        name = {"float32": "liboracle_lbm_f32.so", "float64": "liboracle_lbm_f64.so"}[key]
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
        _libs[key] = C.CDLL(path)
    return _libs[key]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _real(dtype):
    return C.c_float if np.dtype(dtype) == np.float32 else C.c_double


def moment_matrix() -> np.ndarray:
    out = np.zeros((19, 19), np.int32)
    lib = _lib(np.float32)
    lib.orc_moment_matrix(_p(out))
    return out


def minv(dtype=np.float64) -> np.ndarray:
    """numpy.linalg.inv(M) (flow_solver_upgrade.py:123)."""
    return np.linalg.inv(moment_matrix().astype(np.float64)).astype(dtype)


def macro_collide(kind, f, solid, om1=1.0, om2=1.0, S=None, cs_dx2=0.0, force=None):
    """One fused macro+collide pass (lbm_kernels.cu:258-436,1018-1117).
    kind: 'bgk' | 'trt' | 'mrt' | 'mrt_smag'. ``force`` (3 values, lattice units)
    adds the Guo body force (not in the reference). Returns f_post, rho, u, v, w."""
    dt = f.dtype
    R = _real(dt)
    f = np.ascontiguousarray(f.reshape(19, -1))
    N = f.shape[1]
    solid = np.ascontiguousarray(solid.reshape(-1).astype(np.uint8))
    f_post = np.empty_like(f)
    rho, u, v, w = (np.empty(N, dt) for _ in range(4))
    S_arr = np.zeros(19, dt) if S is None else np.ascontiguousarray(np.asarray(S, dt))
    Mi = np.ascontiguousarray(minv(dt))
    k = {"bgk": 0, "trt": 1, "mrt": 2, "mrt_smag": 3}[kind]
    F = None if force is None else np.ascontiguousarray(np.asarray(force, dt).reshape(3))
    _lib(dt).orc_macro_collide(C.c_int(k), _p(f), _p(f_post), _p(rho), _p(u), _p(v), _p(w), _p(solid), R(om1),
                               R(om2), _p(S_arr), _p(Mi), R(cs_dx2), None if F is None else _p(F), C.c_int64(N))
    return f_post, rho, u, v, w


def stream_pull(f_in, nx, ny, nz, px, py, pz):
    """lbm_kernels.cu:447-496."""
    f_in = np.ascontiguousarray(f_in.reshape(19, -1))
    out = np.empty_like(f_in)
    _lib(f_in.dtype).orc_stream_pull(_p(f_in), _p(out), C.c_int(nx), C.c_int(ny), C.c_int(nz), C.c_int(int(px)),
                                     C.c_int(int(py)), C.c_int(int(pz)))
    return out


def sponge(f, rho, u, v, w, sigma):
    """lbm_kernels.cu:505-528 (in place on a copy)."""
    f = np.ascontiguousarray(f.reshape(19, -1)).copy()
    dt = f.dtype
    args = [np.ascontiguousarray(a.reshape(-1).astype(dt)) for a in (rho, u, v, w, sigma)]
    _lib(dt).orc_sponge(_p(f), *[_p(a) for a in args], C.c_int64(f.shape[1]))
    return f


def ibb(f_post, f_stream, links):
    """lbm_kernels.cu:542-588 (returns a modified copy of f_stream)."""
    dt = f_post.dtype
    R = _real(dt)
    fp = np.ascontiguousarray(f_post.reshape(19, -1))
    fs = np.ascontiguousarray(f_stream.reshape(19, -1)).copy()
    n = int(links["fluid_flat"].size)
    if n == 0:
        return fs
    a = {k: np.ascontiguousarray(links[k]) for k in links}
    i32 = lambda x: np.ascontiguousarray(x.astype(np.int32))  # noqa: E731
    _lib(dt).orc_ibb(_p(fp), _p(fs), _p(i32(a["fluid_flat"])), _p(i32(a["dir_i"])), _p(i32(OPP[a["dir_i"]])),
                     _p(a["w1"].astype(dt)), _p(a["w2"].astype(dt)), _p(a["q_ge_half"].astype(np.uint8)),
                     _p(i32(a["ff_flat"])), _p(a["ff_valid"].astype(np.uint8)), C.c_int64(n),
                     C.c_int64(fp.shape[1]))
    return fs


def reset_solid(f, solid):
    """lbm_kernels.cu:597-609 (copy)."""
    f = np.ascontiguousarray(f.reshape(19, -1)).copy()
    s = np.ascontiguousarray(solid.reshape(-1).astype(np.uint8))
    _lib(f.dtype).orc_reset_solid(_p(f), _p(s), C.c_int64(f.shape[1]))
    return f


def zou_he(f, solid, nx, ny, nz, axis, plane, interior, bc_type, u_in, rho_out, sign):
    """lbm_kernels.cu:791-969 (copy)."""
    f = np.ascontiguousarray(f.reshape(19, -1)).copy()
    R = _real(f.dtype)
    s = np.ascontiguousarray(solid.reshape(-1).astype(np.uint8))
    _lib(f.dtype).orc_zou_he(_p(f), _p(s), *[C.c_int(int(v)) for v in (nx, ny, nz, axis, plane, interior, bc_type)],
                             R(u_in), R(rho_out), C.c_int(int(sign)))
    return f


def convective(f, solid, nx, ny, nz, axis, outlet, interior, unknown_cax, rho_out):
    """lbm_kernels.cu:1170-1283 (copy)."""
    f = np.ascontiguousarray(f.reshape(19, -1)).copy()
    R = _real(f.dtype)
    s = np.ascontiguousarray(solid.reshape(-1).astype(np.uint8))
    _lib(f.dtype).orc_convective(_p(f), _p(s), *[C.c_int(int(v)) for v in (nx, ny, nz, axis, outlet, interior,
                                                                            unknown_cax)], R(rho_out))
    return f


def feq(rho, u, v, w, dtype=np.float32):
    """Equilibrium of compute_feq (lbm_kernels.cu:77-113) for arrays of nodes."""
    rho, u, v, w = (np.asarray(a, dtype) for a in (rho, u, v, w))
    t = dtype(1.5) * (u * u + v * v + w * w)
    ox, oy, oz = dtype(3) * u, dtype(3) * v, dtype(3) * w
    out = np.empty((19,) + rho.shape, dtype)
    for q in range(19):
        cu = np.zeros_like(rho)
        for c, o in ((CX[q], ox), (CY[q], oy), (CZ[q], oz)):
            if c:
                cu = cu + (o if c > 0 else -o)
        out[q] = W32[q].astype(dtype) * rho * (dtype(1) + cu + dtype(0.5) * cu * cu - t)
    return out


def step(f, solid_flow, nx, ny, nz, periodic, kind="bgk", om1=1.0, om2=1.0, S=None, cs_dx2=0.0, links=None,
         sigma=None, bcs=(), force=None):
    """One reference step (lbm_cuda_ext.py:445-517): collide -> sponge -> stream_pull
    -> IBB + reset_solid (or bounce-back at solid nodes) -> flow BCs.
    Returns f_next and the pre-collision macros (rho, u, v, w) with u = 0 in solids.
    ``force``: Guo body force in the collision (not in the reference)."""
    fp, rho, u, v, w = macro_collide(kind, f, solid_flow, om1, om2, S, cs_dx2, force)
    sf = solid_flow.reshape(-1).astype(bool)
    u[sf] = 0; v[sf] = 0; w[sf] = 0
    if sigma is not None:
        fp = sponge(fp, rho, u, v, w, sigma)
    fs = stream_pull(fp, nx, ny, nz, *periodic)
    if links is not None and int(links["fluid_flat"].size) > 0:
        fs = ibb(fp, fs, links)
        fs = reset_solid(fs, solid_flow)
    else:  # _bounce_back_post_stream (LBM_CHT.py:1330-1336)
        fc = fs.copy()
        for q in range(19):
            fs[q, sf] = fc[OPP[q], sf]
    for bc in bcs:
        if bc["kind"] == "zou_he":
            fs = zou_he(fs, solid_flow, nx, ny, nz, bc["axis"], bc["plane"], bc["interior"], bc["bc_type"],
                        bc["u_in"], bc["rho_out"], bc["sign"])
        else:
            fs = convective(fs, solid_flow, nx, ny, nz, bc["axis"], bc["plane"], bc["interior"], bc["unknown_cax"],
                            bc["rho_out"])
    return fs, rho, u, v, w


# --------------------------------------------------------------------------
# Bouzidi links (pre_post/sdf_tracing.py:20-85): for q = 1..18 in order, the
# fluid nodes (C order) whose in-range neighbour x + c_q is solid; second node
# x_ff = x - c_q, valid when in range and fluid.
# --------------------------------------------------------------------------
def build_links(solid: np.ndarray, periodic=(False, False, False)):
    """``periodic`` (a build extension; the reference never wraps) also links
    across periodic faces."""
    solid = np.asarray(solid, bool)
    nx, ny, nz = solid.shape
    fluid = ~solid
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ff_all, d_all, fff_all, ok_all = [], [], [], []

    def wrap(a, n, p):
        return a % n if p else a

    for q in range(1, 19):
        cx, cy, cz = C_NP[q]
        inb, jnb, knb = wrap(I + cx, nx, periodic[0]), wrap(J + cy, ny, periodic[1]), wrap(K + cz, nz, periodic[2])
        valid = (inb >= 0) & (inb < nx) & (jnb >= 0) & (jnb < ny) & (knb >= 0) & (knb < nz)
        m = valid & fluid & solid[inb.clip(0, nx - 1), jnb.clip(0, ny - 1), knb.clip(0, nz - 1)]
        if not m.any():
            continue
        i, j, k = I[m], J[m], K[m]
        ff_all.append((i * ny + j) * nz + k)
        d_all.append(np.full(i.size, q, np.int64))
        iff, jff, kff = wrap(i - cx, nx, periodic[0]), wrap(j - cy, ny, periodic[1]), wrap(k - cz, nz, periodic[2])
        ok = (iff >= 0) & (iff < nx) & (jff >= 0) & (jff < ny) & (kff >= 0) & (kff < nz)
        ok &= fluid[iff.clip(0, nx - 1), jff.clip(0, ny - 1), kff.clip(0, nz - 1)]
        fff_all.append((iff.clip(0, nx - 1) * ny + jff.clip(0, ny - 1)) * nz + kff.clip(0, nz - 1))
        ok_all.append(ok)
    if not ff_all:
        z = np.zeros(0, np.int64)
        return {"fluid_flat": z, "dir_i": z, "ff_flat": z, "ff_valid": np.zeros(0, bool)}
    return {"fluid_flat": np.concatenate(ff_all).astype(np.int64), "dir_i": np.concatenate(d_all),
            "ff_flat": np.concatenate(fff_all).astype(np.int64), "ff_valid": np.concatenate(ok_all)}


def bfl_q(phi, origin, pitch, fluid_flat, dir_i, n_coarse=6, n_bisect=10, periodic=(False, False, False)):
    """q, w1, w2, q>=1/2 per link (sdf_tracing.py:127-259), FP32 eager op order.
    ``periodic`` (build extension): samples on those axes are the node's index plus
    the offset t d / pitch, wrapped (the reference clamps every axis)."""
    phi = np.ascontiguousarray(phi, np.float32)
    nx, ny, nz = phi.shape
    n = int(fluid_flat.size)
    q = np.empty(n, np.float32); w1 = np.empty(n, np.float32); w2 = np.empty(n, np.float32)
    ge = np.empty(n, np.uint8)
    org = np.ascontiguousarray(np.asarray(origin, np.float32))
    ff = np.ascontiguousarray(fluid_flat, np.int64)
    di = np.ascontiguousarray(dir_i, np.int64)
    _lib(np.float32).orc_bfl_q_per(_p(phi), C.c_int(nx), C.c_int(ny), C.c_int(nz), _p(org), C.c_float(pitch),
                                   _p(ff), _p(di), C.c_int64(n), C.c_int(n_coarse), C.c_int(n_bisect),
                                   *[C.c_int(int(bool(p))) for p in periodic], _p(q), _p(w1), _p(w2), _p(ge))
    return q, w1, w2, ge.astype(bool)


def sdf_points(phi, origin, pitch, x):
    """sdf_trilinear (sdf_tracing.py:88-125) at points x (..., 3), FP32 eager op order."""
    phi = np.ascontiguousarray(phi, np.float32)
    pts = np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1, 3))
    out = np.empty(pts.shape[0], np.float32)
    org = np.ascontiguousarray(np.asarray(origin, np.float32))
    _lib(np.float32).orc_sdf_points(_p(phi), *[C.c_int(v) for v in phi.shape], _p(org), C.c_float(pitch), _p(pts),
                                    C.c_int64(pts.shape[0]), _p(out))
    return out.reshape(np.asarray(x).shape[:-1])


def q_rays(phi, origin, pitch, x0, d, L, n_coarse=6, n_bisect=10):
    """bfl_q_from_sdf (sdf_tracing.py:127-184) for arbitrary rays, FP32 eager op order."""
    phi = np.ascontiguousarray(phi, np.float32)
    x0 = np.ascontiguousarray(np.asarray(x0, np.float32).reshape(-1, 3))
    d = np.ascontiguousarray(np.asarray(d, np.float32).reshape(-1, 3))
    L = np.ascontiguousarray(np.asarray(L, np.float32).reshape(-1))
    q = np.empty(L.size, np.float32)
    org = np.ascontiguousarray(np.asarray(origin, np.float32))
    _lib(np.float32).orc_q_rays(_p(phi), *[C.c_int(v) for v in phi.shape], _p(org), C.c_float(pitch), _p(x0), _p(d),
                                _p(L), C.c_int64(L.size), C.c_int(n_coarse), C.c_int(n_bisect), _p(q))
    return q


def bfl_weights(q):
    """bfl_weights_from_q (sdf_tracing.py:187-205): (q >= 1/2, w1, w2)."""
    q = np.ascontiguousarray(np.asarray(q, np.float32).reshape(-1))
    ge = np.empty(q.size, np.uint8)
    w1 = np.empty(q.size, np.float32)
    w2 = np.empty(q.size, np.float32)
    _lib(np.float32).orc_bfl_weights(_p(q), C.c_int64(q.size), _p(ge), _p(w1), _p(w2))
    return ge.astype(bool), w1, w2


def set_threads(n: int) -> None:
    """OpenMP thread count of both oracle libraries."""
    for dt in (np.float32, np.float64):
        _lib(dt).orc_set_threads(C.c_int(int(n)))


def linspace01(steps):
    out = np.empty(steps, np.float32)
    _lib(np.float32).orc_linspace01(C.c_int(steps), _p(out))
    return out


def ibb_from_sdf(solid, phi, origin=(0.0, 0.0, 0.0), pitch=1.0, periodic=(False, False, False)):
    """Complete IBBData-equivalent dict for a geometry (links + q + weights)."""
    links = build_links(solid, periodic)
    q, w1, w2, ge = bfl_q(phi, origin, pitch, links["fluid_flat"], links["dir_i"], periodic=periodic)
    links.update(q=q, w1=w1, w2=w2, q_ge_half=ge)
    return links
