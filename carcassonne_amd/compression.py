"""Bond truncation of a corner-side pair.  Re-implements computeProductCompressor
(carcassonne/compression.py:11-46) on device data.

Four alternating-least-squares sweeps: with the current isometry U the generated
contractor forms A (l*r x old*new) such that A vec(X) = L (X U^H (x) conj) R; the new X
solves the normal equations A^T A x = A^T b (b = vec(L R)), then U = polar(X).  The
reference hands the normal equations to SciPy GMRES (relative tolerance 1e-5); here the
minimum-norm least-squares x comes from the eigen-decomposition of A^T A (one-sided Jacobi
on the GPU) with eigenvalues below 1e-10 of the largest dropped, or, above 256 unknowns, a
Cholesky factorization with a 1e-10 relative diagonal shift.  The start is the reference's
random isometry, or (start="svd") a deterministic one, compressorStart.
"""
import numpy as np
import torch

from . import _lib
from .data import HipData
from .device import empty, ptr, stream
from .utils import Join, prependDataContractor

# 0 = L, 1 = U*, 2 = U^H, 3 = U^T, 4 = R: joins L2-U*0, U^H1-R0, U^T1-R1, U*1-U^T0, L3-R2;
# result ((L0,R3), (L1,U^H0))                                              compression.py:12-25
_form = [
    Join(0, 2, 1, 0),
    Join(2, 1, 4, 0),
    Join(3, 1, 4, 1),
    Join(1, 1, 3, 0),
    Join(0, 3, 4, 2),
], [
    [(0, 0), (4, 3)],
    [(0, 1), (2, 0)],
]


def _solve_normal_equations(A, b, cutoff=1e-10):
    """The minimum-norm least-squares x of A x = b: the normal equations A^T A x = A^T b solved with
    the eigen-decomposition of A^T A (one-sided Jacobi on the GPU, eigenvalues below cutoff * largest
    dropped, i.e. singular values of A below 1e-5 of the largest, where the reference's GMRES stops at a
    1e-5 relative residual).
    Systems with more than 256 unknowns use a Cholesky factorization with a 1e-10 relative shift."""
    At = A.join(1, 0)
    G = At.contractWith(A, (1,), (0,))
    rhs = At.contractWith(b, (1,), (0,))
    n = G.shape[0]
    if n <= 256:
        U, S, Vt = G.svd(False)                       # G symmetric PSD: G = U diag(S) U^T
        s = S.toArray()
        keep = int(np.sum(s > cutoff * max(s[0], 1e-300)))
        if keep == 0:
            return HipData.newZeros((n,))
        Uk = U[:, :keep]
        inv = HipData.fromArray(1.0 / s[:keep])
        c = Uk.contractWith(rhs, (0,), (0,)) * inv
        return Uk.contractWith(c, (1,), (0,))
    g = G.tensor()
    tr = float(torch.diagonal(g).sum().item())
    eps = 1e-10 * max(tr, 1e-300) / n
    shift = HipData.newIdentity(n) * eps
    Lm = (G + shift).tensor().clone()
    info = torch.zeros(1, dtype=torch.int32, device=Lm.device)
    work = empty((_lib.query("cc_linalg_work_size", n),))
    W = empty((n, n))
    _lib.call("cc_cholesky_inverse_f64", stream(), n, ptr(Lm), ptr(W), ptr(info), ptr(work))
    if int(info.item()) != 0:
        raise ValueError("normal equations of the compressor are not positive definite")
    Wt = empty((n, n))
    _lib.call("cc_transpose_f64", stream(), n, ptr(W), ptr(Wt))
    t, x = empty((n,)), empty((n,))
    _lib.call("cc_skinny_f64", stream(), n, 1, ptr(W), 1, ptr(rhs.tensor()), ptr(t))
    _lib.call("cc_skinny_f64", stream(), n, 2, ptr(Wt), 1, ptr(t), ptr(x))
    return HipData(x)


def compressorStart(L, new_dimension):
    """A deterministic start for the compressor iteration (the reference draws a random isometry,
    compression.py:35): the new_dimension leading left singular vectors of L matricized at its inward
    ket leg, (old) x (l * old * w), from the eigenvectors of its Gram matrix, each column's sign fixed so
    that its largest entry is positive; columns past the rank are completed with unit vectors and the
    whole made an isometry."""
    old = L.shape[1]
    M = L.join(1, (0, 2, 3))
    G = M.contractWith(M, (1,), (1,))
    U, _, _ = G.svd(False)
    U = U.tensor().contiguous()
    _lib.call("cc_sign_normalize_columns_f64", stream(), old, old, ptr(U))   # the oracle's sign convention
    U = HipData(U)
    if new_dimension <= old:
        return U[:, :new_dimension]
    u = U.toArray()
    return HipData.fromArray(np.hstack([u, np.identity(old)[:, :new_dimension - old]])).unitize()


@prependDataContractor(*_form)
def computeProductCompressor(formMatrix, L, R, new_dimension, start="random"):
    """Returns the compressor (new x old) for bond (L1, L2) = (R0, R1); compression.py:26-45.
    start="random" is the reference's random isometric start, start="svd" compressorStart."""
    if L.shape[1] != L.shape[2]:
        raise ValueError("left inward dimensions do not match (given " + str(L.shape) + ")")
    if R.shape[0] != R.shape[1]:
        raise ValueError("right inward dimensions do not match (given " + str(R.shape) + ")")
    if L.shape[1] != R.shape[1]:
        raise ValueError("left and right shapes are incompatible (given " + str(L.shape) + " and " + str(R.shape) + ")")
    old_dimension = L.shape[1]
    b = L.contractWith(R, (1, 2, 3), (0, 1, 2)).ravel()
    if start == "svd":
        compressor = compressorStart(L, new_dimension)
    else:
        compressor = HipData.newRandom(old_dimension, new_dimension).unitize()
    for _ in range(4):
        A = formMatrix(L, compressor.conj(), compressor.transpose().conj(), compressor.transpose(), R)
        x = _solve_normal_equations(A, b)
        compressor = x.split(old_dimension, new_dimension).unitize()
    return compressor.transpose()


__all__ = ["computeProductCompressor", "compressorStart"]
