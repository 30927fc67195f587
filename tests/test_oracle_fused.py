"""Pins of the oracle's fused-bond sweep step (oracle/fused.py) by brute force and invariants."""
import numpy as np
import pytest

from oracle import fused
from oracle.system import System as OSystem
from oracle.tags import Identity


def test_basis_is_the_dominant_symmetric_column_space():
    """The k basis vectors span the same space as the k leading eigenvectors of Cs Cs^T (Cs the a <-> b
    symmetrized matrix), they are orthonormal, and each is itself symmetric under a <-> b."""
    rng = np.random.default_rng(3)
    nx, na, far, k = 5, 3, 12, 7
    C = rng.standard_normal((nx * na * na, far))
    U, s = fused.fused_basis(C, nx, na, k)
    np.testing.assert_allclose(U.T @ U, np.identity(k), atol=1e-13)
    C3 = C.reshape(nx, na, na, far)
    Cs = (0.5 * (C3 + C3.transpose(0, 2, 1, 3))).reshape(-1, far)
    w, V = np.linalg.eigh(Cs @ Cs.T)
    P_want = V[:, -k:] @ V[:, -k:].T
    np.testing.assert_allclose(U @ U.T, P_want, atol=1e-11)
    U4 = U.reshape(nx, na, na, k)
    np.testing.assert_allclose(U4, U4.transpose(0, 2, 1, 3), atol=1e-12)
    np.testing.assert_allclose(np.sort(s ** 2)[::-1], np.sort(w)[::-1][:far], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("D,chi,seed", [(2, 6, 1), (3, 4, 2)])
def test_step_at_the_rank_bound_keeps_the_normalization(D, chi, seed):
    """At k = far (the corner's rank bound) the fused step keeps every Identity-channel product, so the
    normalization equals the one after contractTowards alone; H stays symmetric; and the bonds come back
    fused at (chi, 1, 1)."""
    from carcassonne_amd import synthetic as syn
    from tests.oracle_bridge import to_oracle
    for direction in range(4):
        c, s, st = syn.make_environment(D, chi, seed=seed)
        oc, os_, op = to_oracle(c, s, syn.operator_center("ising"))
        ref = OSystem(oc, os_, st, op)
        ref.contractTowards(direction)
        oc, os_, op = to_oracle(c, s, syn.operator_center("ising"))
        sys_ = OSystem(oc, os_, st, op)
        fused.contract_and_compress_towards(sys_, direction, chi)
        assert abs(sys_.computeNormalization() / ref.computeNormalization() - 1) < 1e-11
        for t in list(sys_.corners) + list(sys_.sides):
            sh = t[Identity()].shape
            assert sh[0] == chi and sh[1:3] == (1, 1) and sh[3] == chi and sh[4:6] == (1, 1)
        H = sys_.formOperator().matrix_H()
        assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()
