"""Device eigensolvers of the center-site solve.

* GroundStateSolver -- the residual-controlled FP64 solve on the GPU: the
  environment norm matrix N_env (N = N_env (x) I_d) is formed with one MFMA GEMM
  over the identity-channel slabs, factored N_env = L L^T (blocked Cholesky),
  inverted (W = L^-1), and the lowest eigenpair of W H W^T is found by Lanczos
  with full reorthogonalisation (cc_lanczos_f64: native driver, all vectors and
  recurrence scalars on the device).
* relax_over -- relaxOver (carcassonne/utils.py:805-878) on device vectors: the same
  restarted Arnoldi with Krylov size 3, classical Gram-Schmidt, 1e-14 completeness
  cut-off, projected eig with argmin Re, stop rule, dense-vs-matrix-free choice by
  the Multiplier cost rule, final Rayleigh quotient and RelaxFailed test.  N^-1 is
  applied through a Cholesky factor of the formed normalization matrix (the
  reference uses an LU, or GMRES when forming is dearer; both solve the same
  system).
"""
import ctypes

import numpy as np
import torch
from scipy.linalg import eig

from . import _lib
from .data import HipData
from .device import empty, fork, join, ptr, stream, synchronize
from .errors import RelaxFailed


class NormalizationFactor:
    """N_env = L L^T and W = L^-1 for a normalization Stage3Operator (one identity term)."""

    def __init__(self, n_operator, defer=False):
        """defer: allocate only; the caller runs refresh(n_operator) (on a stream of its choice)."""
        if len(n_operator._A) != 1 or n_operator._O[0] is not None:
            raise ValueError("the normalization operator must be a single identity term")
        P, Q = n_operator.P, n_operator.Q
        self.n_env, self.d = P * Q, n_operator.d
        self._shape = (P, Q, n_operator.n_alpha, n_operator.d)
        self._G = empty((P * P, Q * Q))
        self.N_env = empty((self.n_env, self.n_env))
        self.L, self.W = empty(self.N_env.shape), empty(self.N_env.shape)
        self.Lt, self.Wt = empty(self.N_env.shape), empty(self.N_env.shape)
        self._info = torch.zeros(1, dtype=torch.int32, device=self.L.device)
        self._lwork = empty((_lib.query("cc_linalg_work_size", self.n_env),))
        if not defer:
            self.refresh(n_operator)

    @classmethod
    def beside(cls, n_operator):
        """The factor with its kernels (the N_env product and the Cholesky recursion, whose leaves occupy
        one workgroup each) on the side stream, so that the caller's own setup kernels run meanwhile.
        Returns (factor, side stream); the caller joins the side stream before using the factor."""
        f = cls(n_operator, defer=True)    # buffers from the caller's stream
        n_operator._A[0]._t, n_operator._B[0]._t   # the joined slabs too: nothing is allocated on the side stream
        side = fork()
        with torch.cuda.stream(side):
            f.refresh(n_operator)
        return f, side

    def refresh(self, n_operator):
        """Forms N_env, L, W and their transposes from a normalization operator of this factor's shape into
        the factor's own buffers (kernels only, no allocation or host work, so the sequence can be captured
        in a HIP graph and replayed for new environments)."""
        if len(n_operator._A) != 1 or n_operator._O[0] is not None:
            raise ValueError("the normalization operator must be a single identity term")
        if (n_operator.P, n_operator.Q, n_operator.n_alpha, n_operator.d) != self._shape:
            raise ValueError("refresh needs an operator of the factor's shape {}".format(self._shape))
        A, B = n_operator._A[0], n_operator._B[0]
        P, Q, na, _ = self._shape
        # G[(p'p)][(q'q)] = sum_alpha A[alpha][(p'p)] B[alpha][(q'q)]
        _lib.call("cc_gemm_f64", stream(), 1, P * P, Q * Q, na, 1.0, ptr(A._t), 0, 1, P * P, ptr(B._t), 0, Q * Q, 1,
                  0.0, ptr(self._G), 0, Q * Q, 1)
        _lib.call("cc_permute", stream(), _lib.CC_F64, ptr(self._G), ptr(self.N_env), 4, _lib.i64_array([P, Q, P, Q]),
                  _lib.i64_array([P * Q * Q, Q, Q * Q, 1]))
        self.L.copy_(self.N_env)
        _lib.call("cc_cholesky_inverse_f64", stream(), self.n_env, ptr(self.L), ptr(self.W), ptr(self._info),
                  ptr(self._lwork))
        self._checked = False   # the pivot flag is read by check() where the solver synchronizes anyway
        _lib.call("cc_transpose_f64", stream(), self.n_env, ptr(self.L), ptr(self.Lt))
        _lib.call("cc_transpose_f64", stream(), self.n_env, ptr(self.W), ptr(self.Wt))

    def check(self):
        """Raises ValueError when the Cholesky factorization of N_env met a non-positive pivot.  Reads one
        int from the device, so solvers call it after their own final synchronization."""
        if not self._checked:
            bad = int(self._info.item())
            if bad != 0:
                raise ValueError("the normalization matrix is not positive definite (pivot {})".format(bad))
            self._checked = True

    def apply_n(self, x, y):
        """y = (N_env (x) I_d) x for flat device vectors: the normalization multiplier through the formed
        matrix (one skinny product) instead of the stage-3 kernel."""
        _lib.call("cc_skinny_f64", stream(), self.n_env, 0, ptr(self.N_env), self.d, ptr(x), ptr(y))


class GroundStateSolver:
    """Lowest eigenpair of (H, N) for one environment, FP64 on the GPU."""

    def __init__(self, h_operator, n_operator, factor=None, precision="f64"):
        """precision "f32": the H matvecs run on the FP32-operand kernel (FP32 accumulation; vectors,
        orthogonalisation, tridiagonal solve and N stay FP64), the FP32 form of the plain method."""
        if not h_operator.use_fused:
            raise NotImplementedError("the device Lanczos needs an operator with a fused kernel "
                                      "(P = Q in {4, 9, 16, 25, 36}, d = 2)")
        self.H, self.N = h_operator, n_operator
        side = None
        if factor is None:
            factor, side = NormalizationFactor.beside(n_operator)
        self.factor = factor
        self.n = h_operator.n
        self.precision = precision
        try:
            h_operator._ensure_fused()
            self._plan = h_operator.plan_handle(precision)
        finally:
            if side is not None:
                join(side)
        self._work = None
        self._max_iter = 0

    def solve(self, x0, tol_residual=1e-10, tol_energy=1e-13, max_iter=400, check_every=4):
        """x0: flat device tensor (or HipData); returns (x, theta, residual, iterations)."""
        if isinstance(x0, HipData):
            x0 = x0._t.reshape(-1)
        max_iter = int(min(max_iter, self.n))
        if self._work is None or self._max_iter < max_iter:
            self._work = empty((_lib.query("cc_lanczos_work_size", self.n, max_iter),))
            self._max_iter = max_iter
        x = empty((self.n,))
        res = _lib.f64_array([0, 0, 0])
        f = self.factor
        _lib.call("cc_lanczos_f64", stream(), self._plan, f.n_env, f.d, ptr(f.L), ptr(f.Lt), ptr(f.W), ptr(f.Wt),
                  ptr(x0), float(tol_residual), float(tol_energy), max_iter, int(check_every), ptr(self._work),
                  ptr(x), res)
        f.check()
        return x, float(res[0]), float(res[1]), int(res[2])

    def rayleigh_quotient(self, x):
        """(x^T H x) / (x^T N x) in FP64 on the device."""
        hx, nx = torch.empty_like(x), torch.empty_like(x)
        self.H.apply_flat(x, hx)
        self.factor.apply_n(x, nx)
        out = empty((2,))
        _lib.call("cc_dot_f64", stream(), self.n, ptr(x), ptr(hx), ptr(out[0:1]))
        _lib.call("cc_dot_f64", stream(), self.n, ptr(x), ptr(nx), ptr(out[1:2]))
        e, nn = out.tolist()
        return e / nn


# ---------------------------------------------------------------- relaxOver (reference-compatible)
def _dot(x, y):
    out = empty((1,))
    _lib.call("cc_dot_f64", stream(), x.numel(), ptr(x), ptr(y), ptr(out))
    return float(out.item())


def _axpby(a, x, b, y):
    _lib.call("cc_axpby_f64", stream(), x.numel(), float(a), ptr(x), float(b), ptr(y))


class _DenseInverse:
    """x -> N^-1 x for the formed normalization matrix (relaxOver's LU solve, utils.py:814-817).
    A positive definite N goes through W = L^-1 of its Cholesky factor (N^-1 = W^T W).  One that is
    only numerically semi-definite (an environment right after a bandwidth increase) is factored with
    escalating relative diagonal shifts 1e-14, 1e-12, 1e-10.  An indefinite N (environments that are
    Hermitian but not positive, as the reference's random systems) is inverted through its SVD,
    N^-1 = V diag(1/s) U^T, which is the LU's inverse for any invertible N (n <= 256)."""

    def __init__(self, matrix):
        n = matrix.shape[0]
        base = matrix.tensor()
        trace = float(torch.diagonal(base).sum().item())
        work = empty((_lib.query("cc_linalg_work_size", n),))
        info = torch.zeros(1, dtype=torch.int32, device=base.device)
        self.W, self.Ninv, self.n = empty((n, n)), None, n
        for shift in (0.0, 1e-14, 1e-12, 1e-10):
            L = base.clone()
            if shift:
                L.view(-1)[:: n + 1] += shift * abs(trace) / n
            _lib.call("cc_cholesky_inverse_f64", stream(), n, ptr(L), ptr(self.W), ptr(info), ptr(work))
            if int(info.item()) == 0:
                break
        else:
            if n > 256:
                raise ValueError("the normalization matrix is not positive definite (pivot {}); the HIP path "
                                 "factors N > 256 by Cholesky".format(int(info.item())))
            self.Ninv = self._svd_inverse(base.contiguous(), n)
            return
        self.Wt = empty(L.shape)
        _lib.call("cc_transpose_f64", stream(), n, ptr(self.W), ptr(self.Wt))

    @staticmethod
    def _svd_inverse(A, n):
        U, S, Vt = empty((n, n)), empty((n,)), empty((n, n))
        work = empty((_lib.query("cc_svd_work_size", 1, n, n),))
        _lib.call("cc_svd_f64", stream(), 1, n, n, ptr(A), ptr(U), ptr(S), ptr(Vt), ptr(work))
        Sinv, Us = empty((n,)), empty((n, n))
        _lib.call("cc_unary_f64", stream(), 4, n, ptr(S), 0.0, ptr(Sinv))
        _lib.call("cc_scale_rows_cols_f64", stream(), 1, n, n, ptr(U), ptr(Sinv), ptr(Us))      # U diag(1/s)
        out = empty((n, n))
        # N^-1 = Vt^T (U diag(1/s))^T
        _lib.call("cc_gemm_f64", stream(), 1, n, n, n, 1.0, ptr(Vt), 0, 1, n, ptr(Us), 0, 1, n, 0.0, ptr(out), 0, n, 1)
        return out

    def __call__(self, x):
        y = torch.empty_like(x)
        if self.Ninv is not None:
            _lib.call("cc_skinny_f64", stream(), self.n, 0, ptr(self.Ninv), 1, ptr(x), ptr(y))
            return y
        t = torch.empty_like(x)
        _lib.call("cc_skinny_f64", stream(), self.n, 1, ptr(self.W), 1, ptr(x), ptr(t))
        _lib.call("cc_skinny_f64", stream(), self.n, 2, ptr(self.Wt), 1, ptr(t), ptr(y))
        return y


def dense_lowest_eigenpair(H, N):
    """Lowest eigenpair of the pencil (H, N) of formed n x n device matrices (n <= 256), for operators
    without a fused kernel: N = L L^T (Cholesky with inverse W = L^-1, escalating shifts as
    _DenseInverse), M = W H W^T symmetrised, and the one-sided Jacobi SVD of the positive semi-definite
    M + s I with s = |M|_F >= the spectral radius, whose singular values are the eigenvalues of M plus s.
    Returns (x, value) with x = W^T y N-normalised.  This is scipy.linalg.eigh(H, N)[0] of
    minimizeExpectationUsingFullEigensolver (system/_2d.py:498-502)."""
    n = H.shape[0]
    if n > 256:
        raise NotImplementedError("the dense generalized eigensolve handles n <= 256 (here n = {}); "
                                  "larger operators have the fused kernel".format(n))
    inv = _DenseInverse(N)
    W, Wt = inv.W, inv.Wt
    Hm = H.tensor().contiguous()
    T, M = empty((n, n)), empty((n, n))
    _lib.call("cc_gemm_f64", stream(), 1, n, n, n, 1.0, ptr(W), 0, n, 1, ptr(Hm), 0, n, 1, 0.0, ptr(T), 0, n, 1)
    _lib.call("cc_gemm_f64", stream(), 1, n, n, n, 1.0, ptr(T), 0, n, 1, ptr(Wt), 0, n, 1, 0.0, ptr(M), 0, n, 1)
    Mt = empty((n, n))
    _lib.call("cc_transpose_f64", stream(), n, ptr(M), ptr(Mt))
    _lib.call("cc_axpby_f64", stream(), n * n, 0.5, ptr(Mt), 0.5, ptr(M))
    s = float(np.sqrt(max(_dot(M.reshape(-1), M.reshape(-1)), 0.0))) + 1.0
    shift = empty((n, n))
    _lib.call("cc_fill_f64", stream(), n * n, 0.0, ptr(shift))
    _lib.call("cc_set_identity_f64", stream(), n, ptr(shift))
    _lib.call("cc_axpby_f64", stream(), n * n, s, ptr(shift), 1.0, ptr(M))
    U, S, Vt = empty((n, n)), empty((n,)), empty((n, n))
    work = empty((_lib.query("cc_svd_work_size", 1, n, n),))
    _lib.call("cc_svd_f64", stream(), 1, n, n, ptr(M), ptr(U), ptr(S), ptr(Vt), ptr(work))
    y = Vt[n - 1].contiguous()
    x = empty((n,))
    _lib.call("cc_skinny_f64", stream(), n, 2, ptr(Wt), 1, ptr(y), ptr(x))
    value = float(S[n - 1].item()) - s
    return x, value


def relax_over(initial, expectation_multiplier, normalization_multiplier=None, maximum_number_of_multiplications=None,
               tolerance=1e-7, dimension_of_krylov_space=None, stats=None):
    """relaxOver (utils.py:805-878) with the vectors in device memory."""
    shape = initial.shape
    v = initial._t.reshape(-1).clone()
    nv = np.sqrt(_dot(v, v))
    _axpby(0.0, v, 1.0 / nv, v)
    N = v.numel()
    k = 3 if dimension_of_krylov_space is None else dimension_of_krylov_space

    if normalization_multiplier is None:
        inv_N = lambda x: x
    else:
        inv_N = _DenseInverse(normalization_multiplier.formMatrix())
    if expectation_multiplier.isCheaperToFormMatrix(2 * k):
        Hm = expectation_multiplier.formMatrix().tensor()

        def mul_H(x):
            y = torch.empty_like(x)
            _lib.call("cc_skinny_f64", stream(), N, 0, ptr(Hm), 1, ptr(x), ptr(y))
            return y
    else:
        op = getattr(expectation_multiplier, "operator", None)

        def mul_H(x):
            if op is not None:
                y = torch.empty_like(x)
                op.apply_flat(x, y)
                return y
            return expectation_multiplier(HipData(x.reshape(shape))).tensor().reshape(-1)
    multiply = lambda x: inv_N(mul_H(x))
    initial_value = _dot(v, multiply(v))
    mults, last, complete, restarts = 0, None, (k == N), 0
    while True:
        basis = [v]
        mbasis = []
        for i in range(k):
            mbasis.append(multiply(basis[i]))
            if i < k - 1:
                w = mbasis[i].clone()
                coeffs = [_dot(b, mbasis[i]) for b in basis]     # classical Gram-Schmidt, one pass
                for cf, b in zip(coeffs, basis):
                    _axpby(-cf, b, 1.0, w)
                nrm = np.sqrt(max(_dot(w, w), 0.0))
                if nrm <= 1e-14:
                    complete = True
                    break
                _axpby(0.0, w, 1.0 / nrm, w)
                basis.append(w)
        basis = basis[:len(mbasis)]
        mults += k
        restarts += 1
        M = np.array([[_dot(b, mb) for mb in mbasis] for b in basis])
        evals, evecs = eig(M)
        j = int(np.argmin(evals.real))
        lam, vec = evals[j], evecs[:, j]
        stop = complete or (last is not None and abs(last - lam) <= tolerance) or \
            (maximum_number_of_multiplications is not None and mults >= maximum_number_of_multiplications)
        combo = torch.zeros_like(v)
        for cf, b in zip(vec.real, basis):
            _axpby(float(cf), b, 1.0, combo)
        cn = np.sqrt(_dot(combo, combo))
        _axpby(0.0, combo, 1.0 / cn, combo)
        if stop:
            final_value = _dot(combo, multiply(combo))
            if stats is not None:
                stats.update(restarts=restarts, mults=mults + 2, value=final_value)
            denom = abs(final_value) + abs(initial_value)
            if denom > 0 and (final_value - initial_value) / denom > 1 + 1e-7:
                raise RelaxFailed(initial_value, final_value)
            return HipData(combo.reshape(shape))
        v = combo
        last = lam
