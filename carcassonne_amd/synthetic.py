"""Seeded synthetic environments for the benchmark shapes (host-side input generation).

An environment is four corners and four sides, each a dict channel -> array, in
the reference's rank-6 / rank-8 leg layout with the fused bond chi carried on
the ket leg and unit bra / operator legs:

    corner  (chi, 1, 1, chi, 1, 1)
    side    (chi, 1, 1, chi, 1, 1, D, D)
    state   (D, D, D, D, d)

Channel keys are plain tuples so that any tag system can consume them:
``("I",)`` identity, ``("C",)`` complete, ``("T", k)`` the open two-site
channel of operator pair k pointing at the center (the reference's
TwoSiteOperator(k, CENTER) side tag, carcassonne/sparse.py:137-143).  Sides carry
w = 2 + #pairs channels; corners carry I and C.

Construction (all draws from numpy PCG64(seed), entries N(0,1)/sqrt(size)):

* every side slab T[x, y, :, :] is symmetric in (ket, bra) -- antisymmetric for the
  partner channel of an antisymmetric operator half (iY) -- so every stage-3 term
  A (x) B (x) O is symmetric and H is Hermitian;
* the identity channel is positive: corner entries are |N(0,1)|, side slabs are
  Gram matrices G G^T.  Stage 1 and 2 then produce PSD slabs and
  N = sum_alpha A_alpha (x) B_alpha is positive definite;
* the complete and open channels are a multiple of the identity channel plus a
  symmetric N(0,1)/sqrt(size) fluctuation, as in a converged environment where
  they carry an energy (or a local expectation value) times the norm.

The center operator is the 2D transverse-field Ising model
H = -sum ZZ - h sum X (pairs (Z, -Z)) or the spin-1/2 Heisenberg model in Pauli
units H = J sum (XX + YY + ZZ) written over the reals as pairs (X, X),
(iY, -iY), (Z, Z) (iY = [[0, 1], [-1, 0]]).
"""
import numpy as np

X = np.array([[0.0, 1.0], [1.0, 0.0]])
Z = np.array([[1.0, 0.0], [0.0, -1.0]])
IY = np.array([[0.0, 1.0], [-1.0, 0.0]])


def operator_center(model="ising", h=3.0, J=1.0):
    """Returns dict key -> 2x2 matrix with keys ("I",), ("O",) (one-site term) and
    ("TT", k, direction) (two-site halves, as makeSparseOperator lays them out,
    carcassonne/sparse.py:313-333: up/down pair k at directions 3/1, left/right at 0/2)."""
    op = {("I",): np.identity(2)}
    if model == "ising":
        op[("O",)] = -h * X
        pairs = [(Z, -Z)]
    elif model == "heisenberg":
        pairs = [(X, J * X), (IY, -J * IY), (Z, J * Z)]
    else:
        raise ValueError("unknown model " + model)
    for k, (a, b) in enumerate(pairs):
        op[("TT", k, 3)] = a
        op[("TT", k, 1)] = b
        op[("TT", k, 0)] = a
        op[("TT", k, 2)] = b
    return op


def n_pairs(model):
    return 1 if model == "ising" else 3


def pair_is_antisymmetric(model):
    """Per pair k: whether its operator halves are antisymmetric (iY)."""
    return [False] if model == "ising" else [False, True, False]


class _Mixed:
    """standard_normal draws of a base generator plus sigma times the same draws of a second one: the
    environments of one family share the base environment and differ by a seeded perturbation."""

    def __init__(self, seed, perturb_seed, sigma):
        self.a = np.random.Generator(np.random.PCG64(seed))
        self.b = np.random.Generator(np.random.PCG64(perturb_seed)) if perturb_seed is not None else None
        self.sigma = sigma

    def standard_normal(self, shape):
        x = self.a.standard_normal(shape)
        return x if self.b is None else x + self.sigma * self.b.standard_normal(shape)


def make_environment(D, chi, d=2, model="ising", seed=0, energy=-1.0, local=0.3, noise=1.0, perturb_seed=None,
                     sigma=0.0):
    """Returns (corners, sides, state) -- lists of 4 dicts and the state array.  With perturb_seed every
    raw draw of the seed-`seed` environment gets sigma times the same draw of generator `perturb_seed`
    (before the positivity and symmetry constructions, which therefore still hold)."""
    rng = _Mixed(seed, perturb_seed, sigma)
    npairs = n_pairs(model)

    def sym_noise(shape, sign=1.0):
        g = rng.standard_normal(shape) / np.sqrt(np.prod(shape))
        return 0.5 * (g + sign * np.swapaxes(g, -1, -2))

    corners, sides = [], []
    for _ in range(4):
        cI = np.abs(rng.standard_normal((chi, chi))) / np.sqrt(chi * chi)
        cI /= np.linalg.norm(cI) / np.sqrt(chi)
        cC = energy * cI + noise * sym_noise((chi, chi)) * np.linalg.norm(cI) / np.sqrt(chi)
        corners.append({("I",): cI.reshape(chi, 1, 1, chi, 1, 1),
                        ("C",): cC.reshape(chi, 1, 1, chi, 1, 1)})
    for _ in range(4):
        g = rng.standard_normal((chi, chi, D, D)) / np.sqrt(chi * chi * D * D)
        tI = np.einsum("xyab,xycb->xyac", g, g) * chi * D
        tI /= np.linalg.norm(tI) / np.sqrt(chi * D)
        scale = np.linalg.norm(tI) / np.sqrt(chi * chi * D * D)
        side = {("I",): tI, ("C",): energy * tI + noise * scale * sym_noise((chi, chi, D, D)) * D}
        for k, anti in enumerate(pair_is_antisymmetric(model)):
            if anti:   # partner half of an antisymmetric operator: antisymmetric slabs, zero mean
                side[("T", k)] = noise * scale * sym_noise((chi, chi, D, D), -1.0) * D
            else:
                side[("T", k)] = local * tI + noise * scale * sym_noise((chi, chi, D, D)) * D
        sides.append({key: v.reshape(chi, 1, 1, chi, 1, 1, D, D) for key, v in side.items()})
    state = rng.standard_normal((D, D, D, D, d))
    state /= np.linalg.norm(state)
    return corners, sides, state


def make_double_layer_environment(D, c, d=2, seed=0, rank=4):
    """Returns (corners, sides, state) of a seeded double-layer environment in the reference's per-copy
    layout (corner (c, c, 1, c, c, 1), side (c, c, 1, c, c, 1, D, D); bond = (ket, bra, operator)) with the
    Identity channel only: each corner is sum_r m_r (x) m_r and each side sum_r t_r (x) t_r over `rank`
    single-layer tensors m_r (c, c) and t_r (c, c, D) with N(0,1) entries, each tensor scaled to unit
    Frobenius norm.  N is then a sum of norms of single-layer networks -- positive semi-definite, and
    definite for rank > 1 -- H is Hermitian, and the sweep's absorptions and per-copy compressions
    (U (x) U*) keep both properties.  All draws from numpy PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    corners, sides = [], []
    for _ in range(4):
        m = rng.standard_normal((rank, c, c))
        C = np.einsum("rab,rcd->acbd", m, m)
        corners.append({("I",): (C / np.linalg.norm(C)).reshape(c, c, 1, c, c, 1)})
    for _ in range(4):
        t = rng.standard_normal((rank, c, c, D))
        T = np.einsum("rabs,rcdt->acbdst", t, t)
        sides.append({("I",): (T / np.linalg.norm(T)).reshape(c, c, 1, c, c, 1, D, D)})
    state = rng.standard_normal((D, D, D, D, d))
    state /= np.linalg.norm(state)
    return corners, sides, state
