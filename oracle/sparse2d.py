"""Tag-rule lifts of the dense maps (oracle).  Restates carcassonne/tensors/_2d/sparse.py:12-161.

Pinned by: the closed-form walks (identity operator, Z field) and the
2D-in-1D known answers (tests/test_oracle_system.py); a dropped or duplicated
rule changes those energies.
"""
import itertools

import numpy as np

from . import dense
from .tags import (Identity, Complete, OneSiteOperator, TwoSiteOperator,
                   TwoSiteOperatorCompressed, contract_sparse, standard_terms)


def absorb_side_into_corner_from_left(corner, side):
    """absorbSparseSideIntoCornerFromLeft (sparse.py:12-23); handler args are (corner tag, side tag)."""
    f = dense.absorb_side_into_corner_from_left
    rules = standard_terms(f)
    rules.update({
        (TwoSiteOperator, Identity): lambda c, s: (c.matchesSideIdentityOnLeft(), f),
        (Identity, TwoSiteOperator): lambda c, s: (s.matchesCornerIdentityOnRight(), f),
        (TwoSiteOperator, TwoSiteOperator): lambda c, s: (s.matches(c), f),
        (TwoSiteOperatorCompressed, TwoSiteOperatorCompressed): lambda c, s: (s.matches(c), f),
    })
    return contract_sparse(rules, corner, side)


def absorb_side_into_corner_from_right(corner, side):
    """absorbSparseSideIntoCornerFromRight (sparse.py:24-35).  Note the compressed rule
    binds its arguments the other way round (sparse.py:32), restated as written."""
    f = dense.absorb_side_into_corner_from_right
    rules = standard_terms(f)
    rules.update({
        (TwoSiteOperator, Identity): lambda c, s: (c.matchesSideIdentityOnRight(), f),
        (Identity, TwoSiteOperator): lambda c, s: (s.matchesCornerIdentityOnLeft(), f),
        (TwoSiteOperator, TwoSiteOperator): lambda c, s: (c.matches(s), f),
        (TwoSiteOperatorCompressed, TwoSiteOperatorCompressed): lambda c, s: (s.matches(c), f),
    })
    return contract_sparse(rules, corner, side)


def absorb_center_sos_into_side(direction, side, S, op_center, Sc=None):
    """absorbSparseCenterSOSIntoSide (sparse.py:36-58); handler args are (side tag, op tag)."""
    if Sc is None:
        Sc = S.conj()

    def ss(t, _):
        return dense.absorb_center_ss_into_side(direction, t, S, Sc)

    def sos(t, o):
        return dense.absorb_center_sos_into_side(direction, t, S, o, Sc)
    rules = {
        (Identity, Identity): lambda s, c: (Identity(), ss),
        (Complete, Identity): lambda s, c: (Complete(), ss),
        (Identity, OneSiteOperator): lambda s, c: (Complete(), sos),
        (TwoSiteOperator, Identity): lambda s, c: (s.matchesCenterIdentity(), ss),
        (Identity, TwoSiteOperator): lambda s, c: (c.matchesSideIdentityOutward(direction), sos),
        (TwoSiteOperator, TwoSiteOperator): lambda s, c: (s.matchesCenter(direction, c), sos),
        (TwoSiteOperatorCompressed, Identity): lambda s, c: (s, ss),
    }
    return contract_sparse(rules, side, op_center)


def stage1(corner, side):
    """formExpectationStage1 (sparse.py:72-85); handler args are (corner tag, side tag)."""
    f = dense.stage1
    rules = standard_terms(f)
    rules.update({
        (TwoSiteOperator, Identity): lambda c, s: (c.matchesSideIdentityOnRightForStage1(), f),
        (Identity, TwoSiteOperator): lambda c, s: (s.matchesCornerIdentityOnLeftForStage1(), f),
        (TwoSiteOperator, TwoSiteOperator): lambda c, s: (c.matches(s), f),
        (TwoSiteOperatorCompressed, Identity): lambda c, s: (c.matchesSideIdentityOnRightForStage1(), f),
        (Identity, TwoSiteOperatorCompressed): lambda c, s: (s.matchesCornerIdentityOnLeftForStage1(), f),
        (TwoSiteOperatorCompressed, TwoSiteOperatorCompressed): lambda c, s: (c.matches(s), f),
    })
    return contract_sparse(rules, corner, side)


def stage2(right, left):
    """formExpectationStage2 (sparse.py:86-99); handler args are (right tag, left tag)."""
    f = dense.stage2
    rules = standard_terms(f)
    rules.update({
        (TwoSiteOperator, Identity): lambda r, l: (r.matchesStage1IdentityOnLeft(), f),
        (Identity, TwoSiteOperator): lambda r, l: (l.matchesStage1IdentityOnRight(), f),
        (TwoSiteOperator, TwoSiteOperator): lambda r, l: (l.matches(r), f),
        (TwoSiteOperatorCompressed, Identity): lambda r, l: (r.matchesStage1IdentityOnLeft(), f),
        (Identity, TwoSiteOperatorCompressed): lambda r, l: (l.matchesStage1IdentityOnRight(), f),
        (TwoSiteOperatorCompressed, TwoSiteOperatorCompressed): lambda r, l: (l.matches(r), f),
    })
    return contract_sparse(rules, right, left)


_STAGE3_RULES = {
    (Complete, Identity, Identity): lambda x, y, z: True,
    (Identity, Complete, Identity): lambda x, y, z: True,
    (Identity, Identity, OneSiteOperator): lambda x, y, z: True,
    (TwoSiteOperator, Identity, TwoSiteOperator): lambda x, y, z: x.matchesCenterForStage3(0, z),
    (Identity, TwoSiteOperator, TwoSiteOperator): lambda x, y, z: y.matchesCenterForStage3(1, z),
    (TwoSiteOperator, TwoSiteOperator, Identity): lambda x, y, z: x.matchesForStage3(y),
    (TwoSiteOperatorCompressed, TwoSiteOperatorCompressed, Identity): lambda x, y, z: x.matchesForStage3(y),
}


def stage3_terms(s2_0, s2_1, op_center):
    """formExpectationStage3 enumeration (sparse.py:100-127): the list of
    (tag0, tag1, optag, A, B, Op-or-None) terms in the reference's iteration order.
    Op is None when the op tag is Identity (the normalization contractor)."""
    terms = []
    for tags in itertools.product(s2_0, s2_1, op_center):
        rule = _STAGE3_RULES.get(tuple(type(t) for t in tags))
        if rule is None or not rule(*tags):
            continue
        x, y, z = tags
        op = None if z == Identity() else op_center[z]
        terms.append((x, y, z, s2_0[x], s2_1[y], op))
    return terms


class Stage3Operator:
    """The H and N operators of one solve (the reference's pair of Multipliers,
    sparse.py:129-160), with the reference's costs for relaxOver's decisions."""

    def __init__(self, s2_0, s2_1, op_center):
        self.terms = stage3_terms(s2_0, s2_1, op_center)
        d = next(iter(op_center.values())).shape[0]
        self.d = d
        i0, i1 = s2_0[Identity()], s2_1[Identity()]
        self.state_shape = (i0.shape[2], i0.shape[3], i1.shape[2], i1.shape[3], d)
        self.n = int(np.prod(self.state_shape))
        self._joined = {}
        self.AN, self.BN = dense.join_a(i0), dense.join_b(i1)
        self.s2 = (s2_0, s2_1)
        self.eye = np.identity(d)
        self.cost_H = sum(dense.cost_stage3_multiply(self._j(a, 0).shape, self._j(b, 1).shape, d, op is not None)
                          for (_, _, _, a, b, op) in self.terms)
        self.cost_H_form = sum(dense.cost_stage3_matrix(a.shape, b.shape, d) for (_, _, _, a, b, _) in self.terms)
        self.cost_N = dense.cost_stage3_multiply(self.AN.shape, self.BN.shape, d, False)
        self.cost_N_form = dense.cost_stage3_matrix(i0.shape, i1.shape, d)

    def _j(self, t, which):
        key = (id(t), which)
        if key not in self._joined:
            self._joined[key] = (dense.join_a if which == 0 else dense.join_b)(t)
        return self._joined[key]

    def apply_H(self, v):
        v = v.reshape(self.state_shape)
        out = np.zeros(self.state_shape, dtype=np.result_type(v.dtype, *(a.dtype for (_, _, _, a, _, _) in self.terms)) if self.terms else v.dtype)
        for (_, _, _, a, b, op) in self.terms:
            out = out + dense.stage3_apply(self._j(a, 0), self._j(b, 1), v, op)
        return out

    def apply_N(self, v):
        return dense.stage3_apply(self.AN, self.BN, v.reshape(self.state_shape), None)

    def matrix_H(self):
        m = np.zeros((self.n, self.n), dtype=np.result_type(*(a.dtype for (_, _, _, a, _, _) in self.terms)) if self.terms else float)
        for (_, _, _, a, b, op) in self.terms:
            m = m + dense.stage3_matrix(a, b, self.eye if op is None else op)
        return m

    def matrix_N(self):
        return np.kron(dense.normalization_env_matrix(self.AN, self.BN), self.eye)


def form_multipliers(corners, sides, op_center):
    """formExpectationAndNormalizationMultipliers (sparse.py:59-71)."""
    s2_0 = stage2(stage1(corners[0], sides[0]), stage1(corners[1], sides[1]))
    s2_1 = stage2(stage1(corners[2], sides[2]), stage1(corners[3], sides[3]))
    return Stage3Operator(s2_0, s2_1, op_center)
