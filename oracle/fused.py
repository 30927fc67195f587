"""The sweep step on fused environment bonds (oracle; test infrastructure only).

The reference compresses a corner-side bond per copy, U (x) U* on its ket and bra legs
(carcassonne/system/_2d.py:191-228, carcassonne/compression.py:26-45), with U taken from the
Identity channel of the corner and side.  For an environment whose bond legs are fused into one index
(chi, 1, 1) there is no ket/bra pair to act on; the step restated here takes one basis per grown bond
from the corner's Identity channel instead:

    contract_and_compress_towards(system, i, new): contractTowards(i) (carcassonne/system/_2d.py:435-454),
    then for the two bonds it grows (corner i <-> side i, side i <-> corner R(i)) the grown bond
    (x D, D, 1) = (x, a, b) is cut to the k = min(new, far) leading left singular vectors of the corner's
    Identity channel matricized bond x far, after symmetrizing its rows under a <-> b (the exchange of the
    absorbed center ket and bra legs); every channel of the corner and of the side is projected onto it,
    and the bond becomes (k, 1, 1).

Pinned by tests/test_oracle_fused.py: the basis is the projector onto the dominant symmetric column space
(brute force against eigh of the symmetrized Gram matrix); at k = far (the corner's rank bound) the
normalization after the step equals the one after contractTowards alone (the Identity-channel products
are kept exactly) and H stays symmetric.
"""
import numpy as np

from . import linalg
from .tags import Identity, L, R


def fused_basis(Cm, nx, na, k):
    """Cm: (nx na na) x far, rows (x, a, b).  The k leading left singular vectors of its a <-> b
    symmetric part (each with the sign that makes its largest entry positive: the two ends of a side are
    identified by the absorptions that follow, so the sign of a basis vector is not a free gauge; columns
    whose singular value is at most 1e-7 of the largest are zero), and all singular values."""
    far = Cm.shape[1]
    C3 = Cm.reshape(nx, na, na, far)
    Cs = 0.5 * (C3 + C3.transpose(0, 2, 1, 3))
    u, s, _ = np.linalg.svd(Cs.reshape(nx * na * na, far), full_matrices=False)
    u = linalg.sign_normalize_columns(u[:, :k])
    u[:, s[:k] <= 1e-7 * s[0]] = 0.0      # directions below what a Gram-matrix eigensolve resolves
    return u, s


def _corner_basis(C, bond_legs, far_legs, new):
    sh = C.shape
    b = [sh[j] for j in bond_legs]
    D = b[1]
    assert b[2] == 1 and b[0] % D == 0, b
    nx = b[0] // D
    far = int(np.prod([sh[j] for j in far_legs]))
    k = min(new, far)
    Cm = np.transpose(C, tuple(bond_legs) + tuple(far_legs)).reshape(nx * D * D, far)
    return fused_basis(Cm, nx, D, k)[0]


def contract_and_compress_towards(system, i, new):
    system.contractTowards(i)
    Ul = _corner_basis(system.corners[i][Identity()], (3, 4, 5), (0, 1, 2), new)
    Ur = _corner_basis(system.corners[R(i)][Identity()], (0, 1, 2), (3, 4, 5), new)
    kl, kr = Ul.shape[1], Ur.shape[1]

    def corner_l(t):
        sh = t.shape
        return (t.reshape(sh[0] * sh[1] * sh[2], -1) @ Ul).reshape(sh[:3] + (kl, 1, 1))

    def corner_r(t):
        sh = t.shape
        return (Ur.T @ t.reshape(-1, sh[3] * sh[4] * sh[5])).reshape((kr, 1, 1) + sh[3:])

    def side(t):
        sh = t.shape
        fl, fr = sh[0] * sh[1] * sh[2], sh[3] * sh[4] * sh[5]
        m = t.reshape(fl, fr, -1)
        m = np.einsum("fk,fgz,gl->klz", Ul, m, Ur, optimize=True)
        return m.reshape((kl, 1, 1, kr, 1, 1) + sh[6:])

    system.corners[i] = {tag: corner_l(t) for tag, t in system.corners[i].items()}
    system.corners[R(i)] = {tag: corner_r(t) for tag, t in system.corners[R(i)].items()}
    system.sides[i] = {tag: side(t) for tag, t in system.sides[i].items()}
