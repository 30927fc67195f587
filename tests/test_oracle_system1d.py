"""The NumPy restatement of the reference's 1D system (oracle/system1d.py), pinned by its own identities and
by closed forms: the formed expectation matrix equals the matrix-free multiply; a Z field on a product
state gives exact energies; sweeps of the transverse-field Ising and Heisenberg chains reach the exact
energies per site (-(2/pi)(1+l)E(4l/(1+l)^2) at l = 0.01, 1/4 - ln 2 per bond)."""
import numpy as np
import pytest

from oracle.system1d import System1D, expectation_matrix, expectation_multiply, make_mpo

X = np.array([[0.0, 1.0], [1.0, 0.0]])
Z = np.diag([1.0, -1.0])
iY = np.array([[0.0, 1.0], [-1.0, 0.0]])     # i times Pauli Y, real


def test_expectation_matrix_is_the_multiply():
    rng = np.random.default_rng(0)
    R, L = rng.standard_normal((3, 4, 4)), rng.standard_normal((3, 5, 5))
    O = rng.standard_normal((3, 3, 2, 2))
    M = expectation_matrix(R, L, O)
    n = 4 * 5 * 2
    cols = [expectation_multiply(R, L, O, np.eye(n)[i].reshape(4, 5, 2)).ravel() for i in range(n)]
    np.testing.assert_allclose(M, np.array(cols).T, atol=1e-12)


def test_mpo_layout():
    op, rb, rt, lb, lt = make_mpo(np.identity(2), [Z], [(-0.5 * X, X)])
    assert op.shape == (3, 3, 2, 2) and rb == [1, 0, 0] and lb == [0, 0, 1]
    np.testing.assert_array_equal(op[0, 2], Z)
    np.testing.assert_array_equal(op[0, 1], X)
    np.testing.assert_array_equal(op[1, 2], -0.5 * X)
    np.testing.assert_array_equal(op[0, 0], np.identity(2))
    np.testing.assert_array_equal(op[2, 2], np.identity(2))


def _run(system, runs=4, steps=12):
    """Sweeps of `steps` contractions (alternating directions), each followed by a minimization, with the
    bandwidth raised by 2 (at most doubled) between sweeps."""
    system.minimizeExpectation()
    for r in range(runs):
        if r:
            system.increaseBandwidth(0, by=2, do_as_much_as_possible=True)
        for k in range(steps):
            system.contractTowards(k % 2)
            system.minimizeExpectation()
    return system


def test_z_field_product_state():
    """A Z field alone: the ground state is a product state; every contraction adds exactly -1."""
    op, rb, _, lb, _ = make_mpo(np.identity(2), [Z])
    s = System1D(rb, lb, op, np.ones((1, 1, 2)) / np.sqrt(2))
    s.minimizeExpectation()
    assert abs(s.computeExpectation() + 1) < 1e-9
    for k in range(4):
        e0 = s.computeExpectation()
        s.contractTowards(k % 2)
        assert abs(s.computeExpectation() - e0 + 1) < 1e-9
    assert abs(s.computeEstimatedOneSiteExpectation() + 1) < 1e-9


def test_transverse_ising_chain_energy_per_site():
    from scipy.special import ellipe
    lam = 0.01
    exact = -(2 / np.pi) * (1 + lam) * ellipe(4 * lam / (1 + lam) ** 2)
    assert abs(exact + 1.0000250001562545) < 1e-14
    op, rb, _, lb, _ = make_mpo(np.identity(2), [Z], [(-lam * X, X)])
    s = _run(System1D(rb, lb, op, np.ones((1, 1, 2)) / np.sqrt(2)), runs=3, steps=10)
    assert abs(s.computeEstimatedOneSiteExpectation() - exact) < 1e-9
    assert abs(abs(s.computeOneSiteExpectation()) - abs(exact)) < 1e-9


def test_heisenberg_chain_energy_per_bond():
    op, rb, _, lb, _ = make_mpo(np.identity(2), [], [(X, -X), (iY, iY), (Z, Z)])
    s = _run(System1D(rb, lb, op, np.ones((1, 1, 2)) / np.sqrt(2)), runs=6, steps=16)
    assert abs(s.computeEstimatedOneSiteExpectation() / 4 - (0.25 - np.log(2))) < 5e-4   # bandwidth 10
