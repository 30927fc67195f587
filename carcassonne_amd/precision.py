"""The precision form of the solve: FP16 / BF16 stage-3 matvecs with FP32 accumulation inside the
eigensolver, FP64 for everything that fixes the answer.

LowPrecisionOperator -- the stage-3 operator of one solve with every distinct slab channel
    scaled by 1/max|slab| and rounded once to FP16 or BF16 (per-alpha stride padded to 16 bytes);
    the scales fold into each term's d x d site-operator coefficients.  By default the vector is
    carried as a hi + lo pair and the stage-B intermediate is rounded once (split "vector"): at
    D=6, chi=64 that is as accurate as splitting both (the slab rounding dominates: first FP64
    energy 2.36e-7 vs 2.37e-7 off) at half the stage-B matrix-core work, while rounding the vector
    too leaves the refinement stuck near 1e-5.
MixedSolver -- (1) Lanczos on the whitened low-precision operator (the FP64 whitening of the
    exact N) until the Ritz value moves by less than 1e-6 relative between checks: the rounded
    slabs bound the accuracy of its eigenvector (the FP64 energy of that vector is ~1e-7 off at
    D=6, chi=64 however long it runs), so converging further buys nothing; (2) Jacobi-Davidson
    refinement in FP64: each outer
    step applies the FP64 H and N to one new basis vector, solves the small generalized
    eigenproblem of the basis on the device (FP64 Rayleigh-Ritz), and computes the correction
    from a fixed number of CG steps on the projected, shifted low-precision operator.  The
    energy reported is the FP64 Rayleigh quotient of the final Ritz vector.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .device import empty, join, ptr, stream
from .solver import NormalizationFactor

DTYPES = {"f16": _lib.CC_F16, "bf16": _lib.CC_BF16}
# which operands of the low-precision kernel carry hi + lo pairs: True / "both" -- the vector and the
# stage-B intermediate (only the slabs are rounded); "vector" -- the vector only; False / "none" -- none
SPLIT_MODES = {True: 1, "both": 1, "vector": 2, False: 0, "none": 0}


def _lp_channels(operator):
    """Distinct A and B channels of a stage-3 operator in first-use order and its terms (A index, term,
    B index, site operator) sorted by A channel."""
    a_ids, b_ids, a_list, b_list, rows = {}, {}, [], [], []
    for k, (A, B, O) in enumerate(zip(operator._A, operator._B, operator._O)):
        if id(A) not in a_ids:
            a_ids[id(A)] = len(a_list)
            a_list.append(A)
        if id(B) not in b_ids:
            b_ids[id(B)] = len(b_list)
            b_list.append(B)
        rows.append((a_ids[id(A)], k, b_ids[id(B)], O))
    rows.sort(key=lambda r: (r[0], r[1]))
    return a_list, b_list, rows


class LowPrecisionOperator:
    def __init__(self, operator, dtype="bf16", groups=None, split="vector"):
        if not operator.use_fused:
            raise NotImplementedError("the low-precision kernel needs P = Q in {4, 9, 16, 25, 36} and d = 2")
        self.base, self.dtype, self.split = operator, dtype, {1: "both", 2: "vector", 0: "none"}[SPLIT_MODES[split]]
        self.n, self.n_alpha, self.P, self.d = operator.n, operator.n_alpha, operator.P, operator.d
        code = DTYPES[dtype]
        a_list, b_list, rows = _lp_channels(operator)
        # each channel is converted once (read in place from its stage-2 tensor) and kept on the channel,
        # so operators built from the same stage-2 tensors share the rounded slabs
        fresh = [ch for ch in a_list + b_list if dtype not in ch.lowprec]
        if fresh:
            sc = empty((len(fresh), 513))
            slab_elems = _lib.query("cc_stage3_lp_slab_size", self.n_alpha, self.P)
            bufs = []
            for i, ch in enumerate(fresh):
                buf = torch.empty(slab_elems, dtype=torch.float16 if dtype == "f16" else torch.bfloat16, device=sc.device)
                base, (na0, na1, d0, d1), st = ch.source()
                exact = base.numel() == na0 * na1 * d0 * d1 * d0 * d1   # the map covers the stage-2 tensor
                _lib.call("cc_stage3_lp_convert_strided", stream(), code, na0, na1, d0, d1, _lib.i64_array(st),
                          base.numel() if exact else 0, ptr(base), ptr(sc[i]), ptr(buf))
                bufs.append(buf)
            for i, (ch, buf) in enumerate(zip(fresh, bufs)):
                ch.lowprec[dtype] = (buf, sc[i, 0:1])          # the scale stays on the device
        self._slabs = [ch.lowprec[dtype][0] for ch in a_list + b_list]
        self._structure = [(r[0], r[2]) for r in rows], [(ch.n_alpha, ch.P) for ch in a_list + b_list]
        # channel scales gathered on the device; the plan applies them to the unscaled coefficients
        self._scales = torch.cat([ch.lowprec[dtype][1] for ch in a_list + b_list])
        coef = []
        for (a, _, b, O) in rows:
            m = np.identity(self.d) if O is None else O.toArray()
            coef.extend(m.ravel().tolist())
        self.groups = int(groups if groups is not None else max(1, min(self.n_alpha // 8, 2048)))
        self._work = torch.empty(_lib.query("cc_stage3_lp_work_size", self.n_alpha, self.P, self.d, self.groups),
                                 dtype=torch.float32, device=self._slabs[0].device)
        self._ptrs_a = _lib.ptr_array([t.data_ptr() for t in self._slabs[:len(a_list)]])
        self._ptrs_b = _lib.ptr_array([t.data_ptr() for t in self._slabs[len(a_list):]])
        h = ctypes.c_void_p()
        _lib.call("cc_stage3_lp_plan_create", ctypes.byref(h), code, len(a_list), self._ptrs_a, len(b_list), self._ptrs_b,
                  len(rows), _lib.i32_array([r[0] for r in rows]), _lib.i32_array([r[2] for r in rows]),
                  _lib.f32_array(coef), self.n_alpha, self.P, self.d, self.groups, SPLIT_MODES[split], ptr(self._work),
                  stream(), ptr(self._scales))
        self.plan = h
        chunks = ctypes.c_int(0)
        _lib.call("cc_stage3_lp_plan_chunks", h, ctypes.byref(chunks))
        self.chunks = chunks.value
        diag = ctypes.c_int(0)
        _lib.call("cc_stage3_lp_plan_diagonal_form", h, ctypes.byref(diag))
        self.diagonal_form = bool(diag.value)
        self.bytes_per_matvec = len(self._slabs) * self.n_alpha * self.P * self.P * 2 + 2 * self.n * 8

    def apply_flat(self, x, y):
        _lib.call("cc_stage3_lp_plan_apply", stream(), self.plan, ptr(x), ptr(y))

    def refresh(self, operator):
        """Converts the channels of `operator` -- the same structure as the one this plan was built for,
        from new stage-2 tensors -- into this plan's slab buffers and rescales the plan's coefficients with
        the new channel scales (kernels only, no allocation after the first call: capturable in a HIP
        graph and replayed for new environments)."""
        a_list, b_list, rows = _lp_channels(operator)
        structure = [(r[0], r[2]) for r in rows], [(ch.n_alpha, ch.P) for ch in a_list + b_list]
        if structure != self._structure:
            raise ValueError("refresh needs an operator of the structure this plan was built for")
        if getattr(self, "_refresh_sc", None) is None:
            self._refresh_sc = empty((len(self._slabs), 513))
        code = DTYPES[self.dtype]
        for i, (ch, buf) in enumerate(zip(a_list + b_list, self._slabs)):
            base, (na0, na1, d0, d1), st = ch.source()
            exact = base.numel() == na0 * na1 * d0 * d1 * d0 * d1
            _lib.call("cc_stage3_lp_convert_strided", stream(), code, na0, na1, d0, d1, _lib.i64_array(st),
                      base.numel() if exact else 0, ptr(base), ptr(self._refresh_sc[i]), ptr(buf))
        self._scales.copy_(self._refresh_sc[:, 0])
        _lib.call("cc_stage3_lp_plan_rescale", stream(), self.plan, ptr(self._scales))

    def __del__(self):
        if getattr(self, "plan", None) is not None:
            try:
                _lib.load().cc_stage3_lp_plan_destroy(self.plan)
            except Exception:
                pass


def _schedule(inner_iters, k):
    """CG steps of the k-th correction (0-based): a fixed count, or a list whose last entry repeats."""
    if isinstance(inner_iters, (list, tuple)):
        return int(inner_iters[min(k, len(inner_iters) - 1)])
    return int(inner_iters)


class MixedSolver:
    """Lowest eigenpair of (H, N) to FP64 accuracy with low-precision matvecs doing the bulk."""

    def __init__(self, h_operator, n_operator, dtype="bf16", factor=None, groups=None, split="vector"):
        self.H, self.N = h_operator, n_operator
        side = None
        if factor is None:
            factor, side = NormalizationFactor.beside(n_operator)
        self.factor = factor
        try:
            self.Hlp = LowPrecisionOperator(h_operator, dtype, groups=groups, split=split)
            self.n = h_operator.n
            h_operator._ensure_fused()
            self._hplan = h_operator.plan_handle()     # N is applied through the formed N_env (factor.apply_n)
            self._lz_work, self._lz_max = None, 0
            self._jd_work = empty((_lib.query("cc_jd_work_size", self.n),))
        finally:
            if side is not None:
                join(side)

    def refresh(self, h_operator, n_operator):
        """The solve's setup for a new environment of the same structure, into this solver's buffers: the
        normalization factor, the FP64 packed slabs of the H operator this solver was built with, and the
        low-precision slabs and coefficients (kernels only: capturable in a HIP graph)."""
        self.factor.refresh(n_operator)
        self.H.refresh_fused_from(h_operator)
        self.Hlp.refresh(h_operator)

    def solve(self, x0, max_lowprec_iter=300, check_every=4, inner_iters=20, max_outer=8, tol_energy=1e-13,
              tol_predicted=1e-11, lp_tol_residual=0.0, lp_tol_energy=1e-6):
        """Stops after the FP64 Rayleigh-Ritz value rho_k when |rho_k - rho_{k-1}| <= tol_energy |rho_k|, or
        when the remaining error predicted from the last two changes (geometric convergence,
        d_k * q / (1 - q) with q = d_k / d_{k-1}) is at most tol_predicted |rho_k|."""
        from .data import HipData
        if isinstance(x0, HipData):
            x0 = x0._t.reshape(-1)
        f, n = self.factor, self.n
        max_lowprec_iter = int(min(max_lowprec_iter, n))
        if self._lz_work is None or self._lz_max < max_lowprec_iter:
            self._lz_work = empty((_lib.query("cc_lanczos_work_size", n, max_lowprec_iter),))
            self._lz_max = max_lowprec_iter
        x = empty((n,))
        res = _lib.f64_array([0, 0, 0])
        # (1) low-precision Lanczos until the Ritz value stops moving at the precision floor
        _lib.call("cc_lanczos_lp", stream(), self.Hlp.plan, f.n_env, f.d, ptr(f.L), ptr(f.Lt), ptr(f.W), ptr(f.Wt),
                  ptr(x0), float(lp_tol_residual), float(lp_tol_energy), max_lowprec_iter, int(check_every),
                  ptr(self._lz_work), ptr(x), res)
        lp_iters = int(res[2])
        # (2) Jacobi-Davidson in FP64
        kmax = max_outer + 1
        U, HU, NU = empty((kmax, n)), empty((kmax, n)), empty((kmax, n))
        small = empty((kmax + 1,))
        xr, r, t = empty((n,)), empty((n,)), empty((n,))
        U[0].copy_(x)
        k, rho, prev = 0, None, None
        fp64_passes = 0
        history = []
        stopped_by = None
        for outer in range(max_outer):
            # FP64 H and N on the newest basis vector, then the projected pencil
            _lib.call("cc_stage3_plan_apply", stream(), self._hplan, ptr(U[k]), ptr(HU[k]))
            f.apply_n(U[k], NU[k])
            fp64_passes += 1
            k += 1
            hs, ns = empty((k, k)), empty((k, k))
            for j in range(k):
                _lib.call("cc_basis_dots_f64", stream(), k, n, ptr(U), ptr(HU[j]), ptr(hs[:, j]), k)
                _lib.call("cc_basis_dots_f64", stream(), k, n, ptr(U), ptr(NU[j]), ptr(ns[:, j]), k)
            _lib.call("cc_small_geneig_f64", stream(), k, ptr(hs), ptr(ns), ptr(small[0:1]), ptr(small[1:1 + k]))
            rho = float(small[0].item())
            if not np.isfinite(rho):
                raise FloatingPointError("mixed-precision solve: the FP64 Rayleigh-Ritz value is not finite "
                                         "(history {})".format(history))
            history.append(rho)
            coeffs = small[1:1 + k].contiguous()
            _lib.call("cc_basis_combine_f64", stream(), k, n, ptr(U), ptr(coeffs), ptr(xr))
            if prev is not None:
                delta = abs(prev - rho)
                if delta <= tol_energy * abs(rho):
                    stopped_by = "energy change"
                    break
                if len(history) >= 3:
                    q = delta / max(abs(history[-2] - history[-3]), 1e-300)
                    if q < 1.0 and delta * q / (1.0 - q) <= tol_predicted * abs(rho):
                        stopped_by = "predicted error"
                        break
            prev = rho
            if outer == max_outer - 1:
                break
            # residual r = H xr - rho N xr from the basis products
            _lib.call("cc_basis_combine_f64", stream(), k, n, ptr(HU), ptr(coeffs), ptr(r))
            _lib.call("cc_basis_combine_f64", stream(), k, n, ptr(NU), ptr(coeffs), ptr(t))
            _lib.call("cc_axpby_f64", stream(), n, -rho, ptr(t), 1.0, ptr(r))
            _lib.call("cc_jd_correction_lp", stream(), self.Hlp.plan, f.n_env, f.d, ptr(f.Lt), ptr(f.W), ptr(f.Wt),
                      ptr(small[0:1]), ptr(xr), ptr(r), _schedule(inner_iters, outer), ptr(self._jd_work), ptr(t))
            # orthogonalise the correction against the basis (twice, Euclidean) and normalise
            for _ in range(2):
                c = empty((k,))
                _lib.call("cc_basis_dots_f64", stream(), k, n, ptr(U), ptr(t), ptr(c), 1)
                tmp = empty((n,))
                _lib.call("cc_basis_combine_f64", stream(), k, n, ptr(U), ptr(c), ptr(tmp))
                _lib.call("cc_axpby_f64", stream(), n, -1.0, ptr(tmp), 1.0, ptr(t))
            _lib.call("cc_normalize_f64", stream(), n, ptr(t), ptr(U[k]))
        f.check()
        converged = stopped_by is not None
        return xr, rho, dict(lowprec_iterations=lp_iters, outer_steps=k, fp64_passes=fp64_passes, converged=converged,
                             stopped_by=stopped_by or "max_outer",
                             lowprec_ritz=float(res[0]), lowprec_residual=float(res[1]), history=history)

    def solve_fixed(self, x0, lp_iters=60, outer=3, inner_iters=20):
        """The same solve with a fixed amount of work and no host reads or synchronisation (so it can be
        captured in a HIP graph): lp_iters low-precision Lanczos steps, then `outer` FP64 Rayleigh-Ritz
        passes with outer - 1 corrections between them.  Returns (x, rho, record) as device tensors;
        rho[0] is the FP64 Rayleigh-Ritz value of x, record[k] (k < outer) that of pass k, and
        record[outer], record[outer + 1] are |H x - rho N x|^2 and |N x|^2 of the final vector (the
        convergence record that lets a caller check, after a replay, that the fixed work converged)."""
        from .data import HipData
        if isinstance(x0, HipData):
            x0 = x0._t.reshape(-1)
        f, n = self.factor, self.n
        lp_iters = int(min(lp_iters, n))
        if self._lz_work is None or self._lz_max < lp_iters:
            self._lz_work = empty((_lib.query("cc_lanczos_work_size", n, lp_iters),))
            self._lz_max = lp_iters
        x = empty((n,))
        _lib.call("cc_lanczos_lp", stream(), self.Hlp.plan, f.n_env, f.d, ptr(f.L), ptr(f.Lt), ptr(f.W), ptr(f.Wt),
                  ptr(x0), 0.0, 0.0, lp_iters, 0, ptr(self._lz_work), ptr(x), None)
        kmax = outer
        U, HU, NU = empty((kmax, n)), empty((kmax, n)), empty((kmax, n))
        small = empty((kmax + 1,))
        xr, r, t, tmp = empty((n,)), empty((n,)), empty((n,)), empty((n,))
        hist = empty((outer + 2,))
        U[0].copy_(x)
        for k in range(1, outer + 1):
            _lib.call("cc_stage3_plan_apply", stream(), self._hplan, ptr(U[k - 1]), ptr(HU[k - 1]))
            f.apply_n(U[k - 1], NU[k - 1])
            hs, ns = empty((k, k)), empty((k, k))
            for j in range(k):
                _lib.call("cc_basis_dots_f64", stream(), k, n, ptr(U), ptr(HU[j]), ptr(hs[:, j]), k)
                _lib.call("cc_basis_dots_f64", stream(), k, n, ptr(U), ptr(NU[j]), ptr(ns[:, j]), k)
            _lib.call("cc_small_geneig_f64", stream(), k, ptr(hs), ptr(ns), ptr(small[0:1]), ptr(small[1:1 + k]))
            coeffs = small[1:1 + k]
            _lib.call("cc_basis_combine_f64", stream(), k, n, ptr(U), ptr(coeffs), ptr(xr))
            _lib.call("cc_axpby_f64", stream(), 1, 1.0, ptr(small[0:1]), 0.0, ptr(hist[k - 1:k]))
            _lib.call("cc_basis_combine_f64", stream(), k, n, ptr(HU), ptr(coeffs), ptr(r))
            _lib.call("cc_basis_combine_f64", stream(), k, n, ptr(NU), ptr(coeffs), ptr(tmp))
            _lib.call("cc_axpby_dev_f64", stream(), n, ptr(small[0:1]), -1.0, ptr(tmp), 1.0, ptr(r))
            if k == outer:     # the final residual H x - rho N x and N x, for the convergence record
                _lib.call("cc_dot_f64", stream(), n, ptr(r), ptr(r), ptr(hist[outer:outer + 1]))
                _lib.call("cc_dot_f64", stream(), n, ptr(tmp), ptr(tmp), ptr(hist[outer + 1:outer + 2]))
                break
            _lib.call("cc_jd_correction_lp", stream(), self.Hlp.plan, f.n_env, f.d, ptr(f.Lt), ptr(f.W), ptr(f.Wt),
                      ptr(small[0:1]), ptr(xr), ptr(r), _schedule(inner_iters, k - 1), ptr(self._jd_work), ptr(t))
            for _ in range(2):
                c = empty((k,))
                _lib.call("cc_basis_dots_f64", stream(), k, n, ptr(U), ptr(t), ptr(c), 1)
                _lib.call("cc_basis_combine_f64", stream(), k, n, ptr(U), ptr(c), ptr(tmp))
                _lib.call("cc_axpby_f64", stream(), n, -1.0, ptr(tmp), 1.0, ptr(t))
            _lib.call("cc_normalize_f64", stream(), n, ptr(t), ptr(U[k]))
        return xr, small[0:1], hist
