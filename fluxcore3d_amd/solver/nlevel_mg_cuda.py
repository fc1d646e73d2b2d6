"""N-level geometric multigrid for A theta = b, A = I + dt L (drop-in for solver/nlevel_mg_cuda.py).

Same public surface as the reference's ``NLevelGeometricMGSolver``
(nlevel_mg_cuda.py:38-184): ``build_if_needed``, ``solve``, ``L_apply_fine``,
``_vcycle``, ``_Mop`` and the ``_levels`` dictionaries (``x``, ``r``,
``b_buf``, ``Lo``, ``mask_f``, ``inv_diagM``, coefficients) that mg_pcg.py
reads. The hierarchy (coefficients, restricted fields, masks) is built by HIP
kernels; one V-cycle runs per level: two fused sweeps, one fused
residual+restriction, the coarser level, one fused prolongation+correction
(x += 0.75 e, x *= mask) and two fused sweeps; the coarsest level's sweeps run
in a single workgroup launch.

Coefficient storage: where a level's face coefficients are symmetric (aW of a
cell equals aE of its -x neighbour, bit for bit, and so on: conduction with
uniform rho*c_p and no advection), the hierarchy also keeps one face array per
axis and the kernels read those three arrays instead of the six per-cell ones
and diagL. fc_mg_build_faces checks every cell's seven values read back through
the face form against the per-cell arrays; any difference keeps that level on
the per-cell arrays. Both forms give the same bits.
"""
from __future__ import annotations

from typing import Optional, Tuple

import torch

import ctypes as _C

from .. import _lib
from . import mg_cuda_ext  # noqa: F401  (registers the MG entry points)

Tensor = torch.Tensor
SKIP_RUN = 64  # cells per entry of a level's skip table (the kernels index it by flat index >> 6)


class MGCoef(_C.Structure):
    _fields_ = [(n, _C.c_void_p) for n in ("aE", "aW", "aN", "aS", "aU", "aD", "diagL", "invD", "fx", "fy", "fz",
                                           "rows")]


class NLevelGeometricMGSolver:
    def __init__(self, device, dtype, flow_axis: int, flow_sign: int, thermal_dt_scale: float, omega: float = 0.8,
                 n_pre: int = 2, n_post: int = 2, n_coarse: int = 20, min_coarse_cells: int = 4,
                 face_coefficients: bool = True, fused_first_sweeps: bool = True, coarse_mask: str = "all",
                 skip_empty_rows: bool = False):
        # The reference falls through to its PyTorch operator for CPU or float64 input
        # (nlevel_mg_cuda.py:71-79); this solver runs the HIP kernels only.
        if torch.device(device).type != "cuda":
            raise RuntimeError("NLevelGeometricMGSolver runs on the HIP device only (no CPU path)")
        if dtype != torch.float32:
            raise RuntimeError("the multigrid V-cycle runs in float32 (no float64 path)")
        err = _lib.load_error()
        if err is not None:
            raise RuntimeError(f"NLevelGeometricMGSolver needs libfluxcore_hip.so: {err}")
        self.device = torch.device(device)
        self.dtype = dtype
        self.ax = int(flow_axis)
        self.sign = int(flow_sign)
        self.s = float(thermal_dt_scale)
        self.omega = float(omega)
        self.n_pre = int(n_pre)
        self.n_post = int(n_post)
        self.n_coarse = int(n_coarse)
        self.min_coarse_cells = int(min_coarse_cells)
        self.face_coefficients = bool(face_coefficients)
        self.fused_first_sweeps = bool(fused_first_sweeps)
        # "all": a coarse cell is an unknown where all 8 children are (>= 0.999, the
        # reference's rule); "any": where any child is. "any" suits systems whose
        # non-unknown cells carry zero conductance (a solid conducting on its own):
        # thin parts then stay on the coarse levels instead of dropping out.
        if coarse_mask not in ("all", "any"):
            raise ValueError("coarse_mask must be 'all' or 'any'")
        self.coarse_mask = coarse_mask
        # Skip runs of SKIP_RUN consecutive cells (flat index) without unknowns in the
        # stencil kernels. Valid when every cell of such a run is decoupled (its six coefficients and diagL are 0, which the build
        # checks) and the right-hand sides and start vectors handed to the solver vanish
        # there: those cells then stay exactly 0, and skipping them leaves every other
        # value, sum and iteration count bit for bit as they were.
        self.skip_empty_rows = bool(skip_empty_rows)
        self._ready = False
        self._dt_cache: Optional[float] = None
        self._shape_cache: Optional[tuple] = None
        self._levels: list = []
        self._nlevels = 0
        self._use_cuda_ext = True
        n_slots = int(_lib.call("fc_mg_scalar_slots"))
        self._scal = torch.zeros(n_slots, dtype=torch.float64, device=self.device)
        self._partial = None

    # ------------------------------------------------------------------ build
    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def build_if_needed(self, k, rho_cp, u, v, w, mask, dt) -> None:
        shape = tuple(k.shape)
        dt = float(dt)
        if not self._ready or self._shape_cache != shape:
            self._build_hierarchy(k, rho_cp, u, v, w, mask)
            self._dt_cache = None
        if self._dt_cache != dt:
            self._update_inv_diag(dt)
            self._dt_cache = dt

    def _build_hierarchy(self, k, rho_cp, u, v, w, mask) -> None:
        """Levels until min(n//2) <= min_coarse_cells; coarse fields are 2x2x2
        means, coarse unknowns those whose mean mask is >= 0.999
        (nlevel_mg_cuda.py:188-260)."""
        dev = self.device
        f32 = lambda t: t.to(device=dev, dtype=torch.float32).contiguous()  # noqa: E731
        fields = [f32(k), f32(rho_cp), f32(u), f32(v), f32(w)]
        msk = f32(mask)
        self._levels = []
        stream = self._stream()
        while True:
            nx, ny, nz = fields[0].shape
            n_ax = (nx, ny, nz)[self.ax]
            out_idx = n_ax - 1 if self.sign == +1 else 0
            dx = float(2 ** len(self._levels))
            shape = (nx, ny, nz)
            lv = {"shape": shape, "out_idx": out_idx, "mask_f": msk}
            for name in ("aE", "aW", "aN", "aS", "aU", "aD", "diagL", "inv_diagM"):
                lv[name] = torch.empty(shape, device=dev, dtype=torch.float32)
            # zeros: with skipped rows no kernel writes those rows of the work vectors
            for name in ("x", "r", "b_buf", "Lo", "tmp"):
                lv[name] = torch.zeros(shape, device=dev, dtype=torch.float32)
            _lib.call("fc_mg_build_coeffs", *[t.data_ptr() for t in fields], nx, ny, nz, self.ax, out_idx, dx, self.s,
                      *[lv[n].data_ptr() for n in ("aE", "aW", "aN", "aS", "aU", "aD", "diagL")], stream)
            lv["coef_struct"] = MGCoef(*[lv[n].data_ptr() for n in ("aE", "aW", "aN", "aS", "aU", "aD", "diagL",
                                                                    "inv_diagM")])
            lv["coef"] = _C.addressof(lv["coef_struct"])  # pointer to the struct, kept alive by the level
            lv["faces"] = self._build_faces(lv) if self.face_coefficients else None
            if self.skip_empty_rows:
                self._build_rows(lv)
            self._levels.append(lv)
            if min(nx // 2, ny // 2, nz // 2) <= self.min_coarse_cells:
                break
            cshape = (nx // 2, ny // 2, nz // 2)
            nf = []
            for t in fields:
                c = torch.empty(cshape, device=dev, dtype=torch.float32)
                _lib.call("fc_mg_restrict_field", t.data_ptr(), c.data_ptr(), nx, ny, nz, 0, stream)
                nf.append(c)
            cm = torch.empty(cshape, device=dev, dtype=torch.float32)
            _lib.call("fc_mg_restrict_field", msk.data_ptr(), cm.data_ptr(), nx, ny, nz,
                      1 if self.coarse_mask == "all" else 2, stream)
            fields, msk = nf, cm
        self._nlevels = len(self._levels)
        self._shape_cache = tuple(k.shape)
        n0 = k.numel()
        self._partial = torch.empty(3 * int(_lib.call("fc_mg_partials_count", n0)), device=dev, dtype=torch.float64)
        shapes = " -> ".join(str(lv["shape"]) for lv in self._levels)
        print(f"[MG] {self._nlevels}-level hierarchy: {shapes}")

    def _build_faces(self, lv) -> Optional[Tuple[Tensor, Tensor, Tensor]]:
        """Face form of the level's coefficients, or None where it is not exact."""
        nx, ny, nz = lv["shape"]
        faces = tuple(torch.empty(lv["shape"], device=self.device, dtype=torch.float32) for _ in range(3))
        bad = torch.zeros(1, device=self.device, dtype=torch.int32)
        _lib.call("fc_mg_build_faces", lv["coef"], nx, ny, nz, self.ax, lv["out_idx"], *[f.data_ptr() for f in faces],
                  bad.data_ptr(), self._stream())
        if int(bad.item()) != 0:
            return None
        st = lv["coef_struct"]
        st.fx, st.fy, st.fz = (f.data_ptr() for f in faces)
        return faces

    def _build_rows(self, lv) -> None:
        """Skip table of one level: entry s is 1 where cells SKIP_RUN s .. SKIP_RUN s +
        SKIP_RUN - 1 (flat index) hold an unknown (mask != 0)."""
        n = lv["mask_f"].numel()
        pad = (-n) % SKIP_RUN

        def runs_any(t: Tensor) -> Tensor:
            flat = t.reshape(-1)
            if pad:
                flat = torch.cat([flat, flat.new_zeros(pad)])
            return flat.reshape(-1, SKIP_RUN).any(dim=1)

        live = runs_any(lv["mask_f"] != 0)
        dead = ~live
        if bool(dead.any()):
            coupled = torch.zeros_like(live)
            for name in ("aE", "aW", "aN", "aS", "aU", "aD", "diagL"):
                coupled |= runs_any(lv[name] != 0)
            if bool((coupled & dead).any()):
                raise ValueError("skip_empty_rows: a run of cells without unknowns has non-zero coefficients")
        lv["rows"] = live.to(torch.uint8).contiguous()
        lv["coef_struct"].rows = lv["rows"].data_ptr()

    def _update_inv_diag(self, dt: float) -> None:
        for lv in self._levels:
            _lib.call("fc_mg_inv_diag", lv["diagL"].data_ptr(), lv["mask_f"].data_ptr(), float(dt), self.omega,
                      lv["inv_diagM"].data_ptr(), lv["diagL"].numel(), self._stream())
        self._ready = True

    # ------------------------------------------------------------------ kernels
    def _Mop(self, T: Tensor, level: int) -> None:
        """Lo = T + dt L(T) into levels[level]['Lo']."""
        self._Mop_cuda(T, level, self._levels[level]["Lo"])

    def _Mop_cuda(self, T: Tensor, level: int, Lo: Tensor) -> None:
        lv = self._levels[level]
        nx, ny, nz = lv["shape"]
        T = T.contiguous()
        _lib.call("fc_mg_mop_fused", T.data_ptr(), *[lv[n].data_ptr() for n in ("aE", "aW", "aN", "aS", "aU", "aD",
                                                                                  "diagL")],
                  Lo.data_ptr(), float(self._dt_cache), nx, ny, nz, self.ax, lv["out_idx"], self._stream())

    def _sweeps(self, level: int, count: int, zero_guess: bool) -> None:
        """`count` fused Jacobi sweeps on levels[level]['x'] (ping-pong through 'tmp').
        From a zero guess the first two run as one pass that forms x1 = invD b at
        the stencil points (same bits, no x1 written or read)."""
        lv = self._levels[level]
        nx, ny, nz = lv["shape"]
        dt = float(self._dt_cache)
        src, dst = lv["x"], lv["tmp"]
        first = 0
        if zero_guess and count >= 2 and self.fused_first_sweeps:
            _lib.call("fc_mg_sweep2_from_zero", src.data_ptr(), lv["b_buf"].data_ptr(), lv["coef"], dt, nx, ny, nz,
                      self.ax, lv["out_idx"], self._stream())
            first = 2
        for s in range(first, count):
            _lib.call("fc_mg_sweep", src.data_ptr(), dst.data_ptr(), lv["b_buf"].data_ptr(), lv["coef"], dt, nx, ny,
                      nz, self.ax, lv["out_idx"], int(zero_guess and s == 0), self._stream())
            src, dst = dst, src
        if src is not lv["x"]:
            lv["x"], lv["tmp"] = lv["tmp"], lv["x"]

    def _vcycle(self, level: int, x_is_zero: bool = False) -> None:
        """One V(n_pre, n_post) cycle on levels[level] (nlevel_mg_cuda.py:271-307)."""
        lv = self._levels[level]
        nx, ny, nz = lv["shape"]
        dt = float(self._dt_cache)
        stream = self._stream()
        if level == self._nlevels - 1:
            _lib.call("fc_mg_coarse_solve", lv["x"].data_ptr(), lv["tmp"].data_ptr(), lv["b_buf"].data_ptr(),
                      lv["coef"], dt, nx, ny, nz, self.ax, lv["out_idx"], self.n_coarse, int(x_is_zero), stream)
            return
        if x_is_zero and self.n_pre == 0:
            lv["x"].zero_()
        self._sweeps(level, self.n_pre, x_is_zero)
        lc = self._levels[level + 1]
        _lib.call("fc_mg_residual_restrict", lv["x"].data_ptr(), lv["b_buf"].data_ptr(), lv["mask_f"].data_ptr(),
                  lv["coef"], dt, nx, ny, nz, self.ax, lv["out_idx"], lc["b_buf"].data_ptr(), stream)
        self._vcycle(level + 1, x_is_zero=True)
        cx, cy, cz = lc["shape"]
        _lib.call("fc_mg_prolong_correct", lc["x"].data_ptr(), lv["x"].data_ptr(), lv["mask_f"].data_ptr(), nx, ny,
                  nz, cx, cy, cz, 0.75, _lib.ptr(lv.get("rows")), stream)
        self._sweeps(level, self.n_post, False)

    # ------------------------------------------------------------------ API
    def solve(self, b: Tensor, x0: Optional[Tensor] = None, max_cycles: int = 50,
              tol: float = 1e-6) -> Tuple[Tensor, dict]:
        """V-cycle iteration to ||b - A x|| / ||b|| < tol (nlevel_mg_cuda.py:102-162)."""
        assert self._ready, "Call build_if_needed() before solve()."
        lv0 = self._levels[0]
        mask_f = lv0["mask_f"]
        x = lv0["x"]
        if x0 is not None:
            x.copy_(x0)
            x.mul_(mask_f)
        else:
            x.zero_()
        lv0["b_buf"].copy_(b)
        lv0["b_buf"].mul_(mask_f)
        n = b.numel()
        stream = self._stream()
        _lib.call("fc_mg_norm2_f32", b.contiguous().data_ptr(), n, self._partial.data_ptr(), self._scal.data_ptr(), 3,
                  stream)
        b_norm = float(self._scal[3].item()) ** 0.5 + 1e-30
        res_rel = res0_rel = float("inf")
        converged = False
        cyc = 0
        for cyc in range(max_cycles):
            self._vcycle(0, x_is_zero=(cyc == 0 and x0 is None))
            x = lv0["x"]
            self._Mop(x, 0)
            _lib.call("fc_mg_residual_f32", lv0["b_buf"].data_ptr(), lv0["Lo"].data_ptr(), None, lv0["r"].data_ptr(),
                      n, self._partial.data_ptr(), self._scal.data_ptr(), 2, stream)
            res_rel = float(self._scal[2].item()) ** 0.5 / b_norm
            if cyc == 0:
                res0_rel = res_rel
            if not (res_rel < float("inf")):
                converged = False
                break
            if cyc > 2 and res_rel > 10.0 * res0_rel:
                converged = False
                break
            if res_rel < tol:
                converged = True
                break
        result = lv0["x"].clone()
        result.mul_(mask_f)
        diverged = (not converged) and (not (res_rel < float("inf")) or (cyc > 2 and res_rel > 10.0 * res0_rel))
        return result, {"converged": converged, "diverged": diverged, "it": cyc + 1, "res_rel": res_rel,
                        "res0_rel": res0_rel}

    def L_apply_fine(self, T: Tensor) -> Tensor:
        """L(T) at the fine level (nlevel_mg_cuda.py:164-184)."""
        assert self._ready
        Lo = torch.empty_like(T)
        self._Mop_cuda(T, 0, Lo)
        Lo.sub_(T)
        Lo.div_(self._dt_cache)
        return Lo

    def precondition(self, r32: Tensor, out: Optional[Tensor] = None, masked_input: bool = False) -> Tensor:
        """z = V-cycle(r) from a zero guess, masked (mg_pcg.py:178-187).

        With ``masked_input`` (r already zero on masked rows, as the flexible-CG
        residual is) the cycle reads r in place and writes z straight into
        ``out``: the fine level's right-hand side and iterate buffers are
        rebound for the call, and the masked rows stay zero because the
        smoother's inverse diagonal and the prolongation carry the mask."""
        lv0 = self._levels[0]
        shape = lv0["shape"]
        if not masked_input:
            lv0["b_buf"].copy_(r32.view(shape))
            lv0["b_buf"].mul_(lv0["mask_f"])
            self._vcycle(0, x_is_zero=True)
            z = lv0["x"] if out is None else out.view(shape).copy_(lv0["x"])
            z.mul_(lv0["mask_f"])
            return z
        b_saved, x_saved, t_saved = lv0["b_buf"], lv0["x"], lv0["tmp"]
        lv0["b_buf"] = r32.view(shape)
        if out is not None:
            lv0["x"] = out.view(shape)
        try:
            self._vcycle(0, x_is_zero=True)
            z = lv0["x"]
        finally:
            lv0["b_buf"] = b_saved
            if out is not None:
                lv0["x"], lv0["tmp"] = x_saved, t_saved
        if out is not None and z.data_ptr() != out.data_ptr():
            out.view(shape).copy_(z)
            z = out.view(shape)
        return z
