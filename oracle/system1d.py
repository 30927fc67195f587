"""The reference's 1D system in NumPy: test infrastructure only (the second side of the 1D <-> 2D
consistency harness, tests/harness_1d2d.py).

Restates carcassonne/system/_1d.py (System: environments, expectation, contraction, bandwidth increase,
relaxOver minimization, the one-site expectation), the contractors of carcassonne/tensors/_1d.py and the
matrix-product operator of carcassonne/sparse.py:265-287 (makeMPO), with utils.py:241-248
(buildProductTensor, buildTensor) and utils.py:333-361 (computeAbsoluteLimitingLinearCoefficient).

Leg conventions (tensors/_1d.py): environments (operator, state, conjugate state); operator
O[left, right, out, in]; state S[right, left, physical].

Pinned by tests/test_oracle_system1d.py: the formed expectation matrix against the matrix-free multiply
(every unit vector); a product-state Z field (exact energies and the contraction's bookkeeping); the
transverse-field Ising chain reaching -(2/pi)(1+l)E(4l/(1+l)^2) per site and the Heisenberg chain
1/4 - ln 2 per bond, both closed forms.
"""
import copy as _copy
from functools import reduce

import numpy as np
from scipy.linalg import eigvals, svd

from . import linalg, solver
from .tags import Complete, Identity, TwoSiteOperator


def build_tensor(dimensions, components, dtype=float):
    """buildTensor (utils.py:244-248)."""
    t = np.zeros(dimensions, dtype=dtype)
    for index, value in components.items():
        t[index] = value
    return t


def build_product_tensor(*factors):
    """buildProductTensor (utils.py:241-242): the outer product of the factors."""
    return reduce(np.multiply.outer, (np.asarray(f, dtype=float) for f in factors))


def make_mpo(I, Os=(), OOs=()):
    """makeMPO (sparse.py:265-287): the operator tensor, the right boundary and its channel tags, the left
    boundary and its channel tags."""
    size = 2 + len(OOs)
    last = size - 1
    components = {(0, 0): I, (last, last): I}
    for O in Os:
        components[0, last] = components[0, last] + O if (0, last) in components else O
    for id_, OO in enumerate(OOs):
        components[0, id_ + 1] = OO[1]
        components[id_ + 1, last] = OO[0]
    dtype = np.result_type(I, *Os, *[x for OO in OOs for x in OO])
    return (build_tensor((size, size, len(I), len(I)), components, dtype),
            [1] + [0] * (size - 1),
            [Identity()] + [TwoSiteOperator(id_, 2) for id_ in range(len(OOs))] + [Complete()],
            [0] * (size - 1) + [1],
            [Complete()] + [TwoSiteOperator(id_, 2) for id_ in range(len(OOs))] + [Identity()])


# -- the contractors of tensors/_1d.py (Join lists written out as index strings) -------------------------
def absorb_center_oss_into_left(L, O, S, Sc):
    """L'[o0, s0, c0] = sum L[a, b, c] O[o0, a, e, f] S[s0, b, f] Sc[c0, c, e] (tensors/_1d.py:11-24)."""
    return np.einsum("abc,xaef,ybf,zce->xyz", L, O, S, Sc, optimize=True)


def absorb_center_oss_into_right(R, O, S, Sc):
    """R'[o1, s1, c1] = sum R[a, b, c] O[a, o1, e, f] S[b, s1, f] Sc[c, c1, e] (tensors/_1d.py:25-38)."""
    return np.einsum("abc,axef,byf,cze->xyz", R, O, S, Sc, optimize=True)


def absorb_center_ss_into_right(R, S, Sc):
    """R'[s1, c1] = sum R[b, c] S[b, s1, f] Sc[c, c1, f] (tensors/_1d.py:51-62)."""
    return np.einsum("bc,byf,czf->yz", R, S, Sc, optimize=True)


def expectation_multiply(R, L, O, S):
    """Y[r2, l2, o2] = sum R[a, b, r2] L[c, e, l2] O[a, c, o2, f] S[b, e, f] (tensors/_1d.py:75-95, the
    inner contractor)."""
    return np.einsum("abx,cey,aczf,bef->xyz", R, L, O, S, optimize=True)


def expectation_matrix(R, L, O):
    """M[(r2, l2, o2), (r1, l1, o3)] = sum R[a, r1, r2] L[c, l1, l2] O[a, c, o2, o3] (tensors/_1d.py:66-74,
    the outer contractor)."""
    m = np.einsum("aux,cvy,aczw->xyzuvw", R, L, O, optimize=True)
    n = R.shape[2] * L.shape[2] * O.shape[2]
    return m.reshape(n, n)


def absolute_limiting_linear_coefficient(n, multiply_O, multiply_N, multiply_L, multiply_R):
    """computeAbsoluteLimitingLinearCoefficient (utils.py:333-361): from the two dominant (generalized)
    eigenvectors of the transfer matrix, sqrt(tr(Om^H Om) - 2) / sqrt(tr(Nm^H Nm))."""
    matrix = np.array([multiply_O(np.eye(n)[i]) for i in range(n)])
    evals = eigvals(matrix)
    lam = evals[np.argmax(np.abs(evals))]
    t = matrix - lam * np.identity(n)
    ovecs = svd(t @ t)[-1][-2:]
    Om = np.array([[np.dot(ovecs[i].conj(), multiply_O(ovecs[j])) for j in range(2)] for i in range(2)])
    numerator = np.sqrt(np.trace(Om.conj().T @ Om) - 2)
    lv, rv = multiply_L(ovecs), multiply_R(ovecs)
    Nm = np.array([[np.dot(lv[i].conj(), multiply_N(rv[j])) for j in range(2)] for i in range(2)])
    return numerator / np.sqrt(np.trace(Nm.conj().T @ Nm))


class System1D:
    """carcassonne/system/_1d.py System."""

    def __init__(self, right_operator_boundary, left_operator_boundary, operator_center_data, state_center_data=None):
        self.right_operator_boundary = np.array(right_operator_boundary, dtype=float)
        self.left_operator_boundary = np.array(left_operator_boundary, dtype=float)
        self.right_environment = build_product_tensor(right_operator_boundary, [1], [1])
        self.left_environment = build_product_tensor(left_operator_boundary, [1], [1])
        self.operator_center_data = np.asarray(operator_center_data)
        if state_center_data is None:
            state_center_data = np.ones((1, 1, self.operator_center_data.shape[2]))
        self.setStateCenter(np.asarray(state_center_data))
        self.rng = np.random.default_rng(0)
        # the polar factors (normalizeAxis, data/__init__.py:263-301); a harness that compares this system
        # with another implementation can give both sides the same routine, since the singular vectors of
        # zero singular values are a free choice of each SVD
        self.normalize_axis = linalg.normalize_axis

    def __copy__(self):
        other = System1D.__new__(System1D)
        other.__dict__.update(self.__dict__)
        return other

    def setStateCenter(self, state, state_conj=None):
        self.state_center_data = state
        self.state_center_data_conj = state.conj() if state_conj is None else state_conj

    def multiply(self, S):
        return expectation_multiply(self.right_environment, self.left_environment, self.operator_center_data, S)

    def formExpectationMatrix(self):
        return expectation_matrix(self.right_environment, self.left_environment, self.operator_center_data)

    def computeExpectation(self):
        """_1d.py:48-50: S* . (H S), real part."""
        return np.dot(self.state_center_data_conj.ravel(), self.multiply(self.state_center_data).ravel()).real

    def computeEstimatedOneSiteExpectation(self, direction=0):
        """_1d.py:88-93."""
        s = _copy.copy(self)
        e1 = s.computeExpectation()
        s.contractTowards(direction)
        return s.computeExpectation() - e1

    def computeOneSiteExpectation(self):
        """_1d.py:51-87."""
        R_shape = self.right_environment.shape
        N_shape = (self.state_center_data.shape[0],) * 2
        S = self.normalize_axis(self.state_center_data, 1)[0]
        Sc = S.conj()
        O = self.operator_center_data
        return absolute_limiting_linear_coefficient(
            int(np.prod(R_shape)),
            lambda v: absorb_center_oss_into_right(v.reshape(R_shape), O, S, Sc).ravel(),
            lambda v: absorb_center_ss_into_right(v.reshape(N_shape), S, Sc).ravel(),
            lambda vs: np.tensordot(self.left_operator_boundary, vs.reshape((vs.shape[0],) + R_shape), (0, 1)).reshape(vs.shape[0], -1),
            lambda vs: np.tensordot(self.right_operator_boundary, vs.reshape((vs.shape[0],) + R_shape), (0, 1)).reshape(vs.shape[0], -1))

    def contractUnnormalizedTowards(self, direction, S=None):
        """_1d.py:140-148 (direction 0: into the right environment, 1: into the left)."""
        S = self.state_center_data if S is None else S
        if direction == 0:
            self.right_environment = absorb_center_oss_into_right(self.right_environment, self.operator_center_data, S, S.conj())
        elif direction == 1:
            self.left_environment = absorb_center_oss_into_left(self.left_environment, self.operator_center_data, S, S.conj())
        else:
            raise ValueError("Direction must be 0 for right or 1 for left, not {}.".format(direction))

    def contractTowards(self, direction):
        """_1d.py:125-129."""
        to_contract, _, to_absorb = self.normalize_axis(self.state_center_data, 1 - direction)
        self.contractUnnormalizedTowards(direction, to_contract)
        self.setStateCenter(linalg.absorb_matrix_at(self.normalize_axis(self.state_center_data, direction)[0],
                                                    direction, to_absorb))

    def increaseBandwidth(self, direction=0, by=None, to=None, do_as_much_as_possible=False, enlargeners=None):
        """_1d.py:160-163 with base.py:114-160 (a three-leg state: the other axis is 1 - axis)."""
        if direction != 0:
            raise ValueError("Direction for bandwidth increase must be 0, not {}.".format(direction))
        axis, O_axis = 0, 1
        S = self.state_center_data
        d, old = S.shape[-1], S.shape[axis]
        if (by is None) == (to is None):
            raise ValueError("Either 'by' or 'to' must be given, not both.")
        new = old + by if by is not None else to
        assert new >= old
        if new == old:
            return None
        if new > d * old:
            if not do_as_much_as_possible:
                raise ValueError("New dimension must be less than the physical dimension times the old dimension "
                                 "({} > {}*{}).".format(new, d, old))
            new = d * old
        n0 = self.normalize_axis(S, O_axis)[0]
        n1 = self.normalize_axis(S, axis)[0]
        eA, eB = linalg.new_enlargener(old, new, self.rng) if enlargeners is None else enlargeners
        S = linalg.absorb_matrix_at(S, axis, eA)
        n0 = linalg.absorb_matrix_at(n0, O_axis, eB)
        n0, S = self._normalize_and_denormalize(n0, O_axis, axis, S)
        n1 = linalg.absorb_matrix_at(n1, axis, eA)
        S = linalg.absorb_matrix_at(S, O_axis, eB)
        n1, S = self._normalize_and_denormalize(n1, axis, O_axis, S)
        self.setStateCenter(S)
        self.contractUnnormalizedTowards(axis, n0)
        self.contractUnnormalizedTowards(O_axis, n1)
        return eA, eB

    def minimizeExpectationUsingFullEigensolver(self):
        """_1d.py:170-173: the lowest eigenvector of the formed expectation matrix."""
        from scipy.linalg import eigh
        _, V = eigh(self.formExpectationMatrix())
        self.setStateCenter(V[:, 0].reshape(self.state_center_data.shape))

    def _normalize_and_denormalize(self, a, axis_to_norm, axis_to_denorm, other):
        """normalizeAxisAndDenormalize (data/__init__.py:302-310)."""
        iso, _, den = self.normalize_axis(a, axis_to_norm)
        return iso, linalg.absorb_matrix_at(other, axis_to_denorm, den)

    def minimizeExpectation(self):
        """_1d.py:164-169: relaxOver on the expectation multiplier alone (a standard eigenproblem), at most
        100 multiplications."""
        shape = self.state_center_data.shape
        v = solver.relax_over(self.state_center_data.ravel(), lambda x: self.multiply(x.reshape(shape)).ravel(),
                              max_mults=100)
        self.setStateCenter(v.reshape(shape))
