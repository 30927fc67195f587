"""Oracle small linear algebra: normalizeAxis identities (reference test_data.py:98-133), the
polar factor, enlargeners, the compressor's generated contractor and the product compressor's
exactness (test_compression.py:9-30)."""
import itertools

import numpy as np
import pytest

from oracle import linalg

RNG = np.random.default_rng(5)


@pytest.mark.parametrize("cplx", [False, True])
def test_normalize_axis_identities(cplx):
    """Reference test_data.py:98-133: iso absorbing den^T gives back the tensor, iso is an
    isometry over the other legs, den nrm^T = I (both variants)."""
    a = RNG.standard_normal((3, 4, 2, 5))
    if cplx:
        a = a + 1j * RNG.standard_normal(a.shape)
    for axis in range(4):
        iso, nrm, den = linalg.normalize_axis(a, axis)
        m = linalg.join(iso, [i for i in range(4) if i != axis], axis)
        np.testing.assert_allclose(m.conj().T @ m, np.eye(a.shape[axis]), atol=1e-12)
        np.testing.assert_allclose(linalg.absorb_matrix_at(iso, axis, den.T), a, atol=1e-12)
        np.testing.assert_allclose(den @ nrm.T, np.eye(a.shape[axis]), atol=1e-10)
        n2, d2 = linalg.normalize_axis(a, axis, True)
        np.testing.assert_allclose(d2 @ n2.T, np.eye(a.shape[axis]), atol=1e-10)


def test_normalize_axis_rejects_too_few_degrees_of_freedom():
    with pytest.raises(ValueError):
        linalg.normalize_axis(RNG.standard_normal((2, 5)), 1)


def test_unitize_is_the_polar_factor():
    m = RNG.standard_normal((6, 4))
    u = linalg.unitize(m)
    np.testing.assert_allclose(u.T @ u, np.eye(4), atol=1e-12)
    p = u.T @ m                                      # m = u p with p symmetric positive definite
    np.testing.assert_allclose(p, p.T, atol=1e-12)
    assert np.linalg.eigvalsh(p).min() > 0


def test_enlargener_is_an_isometry():
    e, ec = linalg.new_enlargener(3, 7, RNG)
    np.testing.assert_allclose(e.T @ e, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(ec, e.conj())
    i, _ = linalg.new_enlargener(4, 4, RNG)
    np.testing.assert_allclose(i, np.eye(4))


def test_compressor_matrix_explicit_loops():
    """compression.py:12-25: A[(l,r),(a,n)] = sum L[l,a,p,w] U*[p,m] U^H[n,b] U^T[m,c] R[b,c,w,r]."""
    l, old, new, w, r = 2, 3, 2, 2, 2
    Lt, Rt = RNG.standard_normal((l, old, old, w)), RNG.standard_normal((old, old, w, r))
    U = linalg.unitize(RNG.standard_normal((old, new)) + 1j * RNG.standard_normal((old, new)))
    want = np.zeros((l, r, old, new), dtype=complex)
    for li, ri, a, n in itertools.product(range(l), range(r), range(old), range(new)):
        want[li, ri, a, n] = sum(Lt[li, a, p, wi] * U.conj()[p, m] * U.conj().T[n, b] * U.T[m, c] * Rt[b, c, wi, ri]
                                 for p, m, b, c, wi in itertools.product(range(old), range(new), range(old), range(old), range(w)))
    np.testing.assert_allclose(linalg.compressor_matrix(Lt, U, Rt), want.reshape(l * r, old * new), atol=1e-12)


def test_compressor_matrix_reproduces_the_product_at_full_rank():
    """With a square real orthogonal U the projector U U^T is the identity, so A vec(U) = vec(L R)."""
    l, old, w, r = 3, 4, 2, 3
    Lt, Rt = RNG.standard_normal((l, old, old, w)), RNG.standard_normal((old, old, w, r))
    U = linalg.unitize(RNG.standard_normal((old, old)))
    b = np.tensordot(Lt, Rt, ((1, 2, 3), (0, 1, 2))).ravel()
    np.testing.assert_allclose(linalg.compressor_matrix(Lt, U, Rt) @ U.ravel(), b, atol=1e-12)


def test_product_compressor_exact_when_rank_suffices():
    """L = L0 E^T, R = E R0 with an isometric enlargener E (old -> 2 old): compressing back to
    `old` reproduces L R exactly (reference test_compression.py:9-30)."""
    l, k, w, r = 3, 3, 2, 3
    L0, R0 = RNG.standard_normal((l, k, k, w)), RNG.standard_normal((k, k, w, r))
    E, _ = linalg.new_enlargener(k, 2 * k, RNG)
    Lt = np.einsum("lpqw,ap,bq->labw", L0, E, E)
    Rt = np.einsum("pqwr,ap,bq->abwr", R0, E, E)
    Ut = linalg.product_compressor(Lt, Rt, k, RNG)
    U = Ut.T
    P = U @ U.T
    full = np.tensordot(Lt, Rt, ((1, 2, 3), (0, 1, 2)))
    comp = np.einsum("lpqw,pa,qb,abwr->lr", Lt, P, P, Rt)
    np.testing.assert_allclose(comp, full, atol=1e-10 * np.abs(full).max())


def test_sign_normalize_columns_rule():
    """Each column is negated exactly when its entry of largest magnitude (the first of equal ones) is
    negative; applying the rule twice changes nothing."""
    from oracle.linalg import sign_normalize_columns
    U = np.array([[0.1, -0.5, 0.3], [-0.9, 0.5, -0.3], [0.2, 0.1, 0.0]])
    V = sign_normalize_columns(U)
    np.testing.assert_array_equal(V[:, 0], -U[:, 0])     # -0.9 is the largest magnitude
    np.testing.assert_array_equal(V[:, 1], -U[:, 1])     # tie -0.5 / 0.5: the first row decides
    np.testing.assert_array_equal(V[:, 2], U[:, 2])      # tie 0.3 / -0.3: the first row is positive
    np.testing.assert_array_equal(sign_normalize_columns(V), V)


def test_min_norm_solution_is_the_truncated_pseudo_inverse_solution():
    """min_norm_solution against the SVD pseudo-inverse (a different method): equal to np.linalg.solve for
    a well-conditioned square system; for A with singular values 1, 1e-3, 1e-7 and a null space the
    normal-equation cut-off 1e-10 keeps exactly the singular values above 1e-5 of the largest."""
    from oracle.linalg import min_norm_solution
    rng = np.random.default_rng(3)
    A = rng.standard_normal((6, 6)) + 6 * np.identity(6)
    b = rng.standard_normal(6)
    np.testing.assert_allclose(min_norm_solution(A, b), np.linalg.solve(A, b), rtol=1e-12, atol=1e-12)
    U, _ = np.linalg.qr(rng.standard_normal((8, 8)))
    V, _ = np.linalg.qr(rng.standard_normal((5, 5)))
    s = np.array([1.0, 1e-3, 1e-7, 0.0, 0.0])
    A = (U[:, :5] * s) @ V.T
    b = rng.standard_normal(8)
    want = np.linalg.pinv(A, rcond=1e-5) @ b
    got = min_norm_solution(A, b)
    np.testing.assert_allclose(got, want, rtol=1e-8, atol=1e-10 * np.linalg.norm(want))
    assert np.linalg.norm(V[:, 2:] .T @ got) < 1e-8 * np.linalg.norm(got)   # nothing along the cut directions


def test_compressor_start_spans_the_dominant_ket_subspace():
    """compressor_start(Lt, k): orthonormal columns spanning the k-dimensional dominant subspace of Lt
    matricized at its inward ket leg (against eigh of the Gram matrix, a different method), the sign rule
    on every column, and the same start for -Lt (signs are fixed by the rule, not by the input)."""
    from oracle.linalg import compressor_start
    rng = np.random.default_rng(5)
    Lt = rng.standard_normal((3, 6, 6, 2))
    k = 4
    U = compressor_start(Lt, k)
    np.testing.assert_allclose(U.T @ U, np.identity(k), atol=1e-12)
    m = np.moveaxis(Lt, 1, 0).reshape(6, -1)
    w, E = np.linalg.eigh(m @ m.T)
    Pk = E[:, -k:] @ E[:, -k:].T
    np.testing.assert_allclose(U @ U.T, Pk, atol=1e-10)
    for j in range(k):
        i = int(np.argmax(np.abs(U[:, j])))
        assert U[i, j] > 0
    np.testing.assert_allclose(compressor_start(-Lt, k), U, atol=1e-12)
