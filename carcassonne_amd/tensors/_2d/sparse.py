"""Tag-rule lifts of the dense contractors.  Re-implements carcassonne/tensors/_2d/sparse.py:12-161:
the same rule tables (which tag pairs combine and the result tag), applied by
carcassonne_amd.sparse.contractSparseTensors.  Stage 3 gathers every compatible
(stage2_0 tag, stage2_1 tag, operator tag) triple into one Stage3Operator, so a
matvec of the whole expectation multiplier is a single HIP call.
"""
import itertools

from .dense import *  # noqa: F401,F403
from .dense import (absorbDenseCenterSOSIntoSide, absorbDenseCenterSOSIntoSideProjected, absorbDenseCenterSSIntoSide,
                    absorbDenseSideIntoCornerFromLeft,
                    absorbDenseSideIntoCornerFromRight, formNormalizationStage1, formNormalizationStage2)
from ...sparse import (Complete, Identity, OneSiteOperator, TwoSiteOperator, TwoSiteOperatorCompressed,
                       addStandardCompleteAndIdentityTerms, contractSparseTensors, getInformationFromOperatorCenter)
from ... import stage3


def absorbSparseSideIntoCornerFromLeft(corner, side):
    """sparse.py:12-23 (handler arguments: corner tag, side tag)."""
    terms = {}
    f = absorbDenseSideIntoCornerFromLeft
    addStandardCompleteAndIdentityTerms(terms, f)
    terms.update({
        (TwoSiteOperator, Identity): lambda c, s: (c.matchesSideIdentityOnLeft(), f),
        (Identity, TwoSiteOperator): lambda c, s: (s.matchesCornerIdentityOnRight(), f),
        (TwoSiteOperator, TwoSiteOperator): lambda c, s: (s.matches(c), f),
        (TwoSiteOperatorCompressed, TwoSiteOperatorCompressed): lambda c, s: (s.matches(c), f),
    })
    return contractSparseTensors(terms, corner, side)


def absorbSparseSideIntoCornerFromRight(corner, side):
    """sparse.py:24-35 (the compressed rule binds side.matches(corner), as written there)."""
    terms = {}
    f = absorbDenseSideIntoCornerFromRight
    addStandardCompleteAndIdentityTerms(terms, f)
    terms.update({
        (TwoSiteOperator, Identity): lambda c, s: (c.matchesSideIdentityOnRight(), f),
        (Identity, TwoSiteOperator): lambda c, s: (s.matchesCornerIdentityOnLeft(), f),
        (TwoSiteOperator, TwoSiteOperator): lambda c, s: (c.matches(s), f),
        (TwoSiteOperatorCompressed, TwoSiteOperatorCompressed): lambda c, s: (s.matches(c), f),
    })
    return contractSparseTensors(terms, corner, side)


def _center_into_side_terms(direction, ss, sos):
    """The rule table of absorbSparseCenterSOSIntoSide (sparse.py:36-58): (side tag, operator tag) ->
    (result tag, dense contractor) with ss for the identity-operator pairs and sos for the others."""
    return {
        (Identity, Identity): lambda s, c: (Identity(), ss),
        (Complete, Identity): lambda s, c: (Complete(), ss),
        (Identity, OneSiteOperator): lambda s, c: (Complete(), sos),
        (TwoSiteOperator, Identity): lambda s, c: (s.matchesCenterIdentity(), ss),
        (Identity, TwoSiteOperator): lambda s, c: (c.matchesSideIdentityOutward(direction), sos),
        (TwoSiteOperator, TwoSiteOperator): lambda s, c: (s.matchesCenter(direction, c), sos),
        (TwoSiteOperatorCompressed, Identity): lambda s, c: (s, ss),
    }


def absorbSparseCenterSOSIntoSide(direction, side, state_center_data, operator_center_data, state_center_data_conj=None):
    """sparse.py:36-58 (handler arguments: side tag, operator tag)."""
    if state_center_data_conj is None:
        state_center_data_conj = state_center_data.conj()

    def ss(side_data, _):
        return absorbDenseCenterSSIntoSide(direction, side_data, state_center_data, state_center_data_conj)

    def sos(side_data, op_data):
        return absorbDenseCenterSOSIntoSide(direction, side_data, state_center_data, op_data, state_center_data_conj)
    return contractSparseTensors(_center_into_side_terms(direction, ss, sos), side, operator_center_data)


def absorbSparseCenterSOSIntoSideProjected(direction, side, state_center_data, operator_center_data, left, right,
                                           state_center_data_conj=None):
    """absorbSparseCenterSOSIntoSide with every result channel projected onto the fused truncation bases
    `left` (the bond to corner `direction`) and `right` (the bond to corner R(direction)) as it is formed
    (absorbDenseCenterSOSIntoSideProjected): the same rule table, no grown side in memory."""
    if state_center_data_conj is None:
        state_center_data_conj = state_center_data.conj()

    def ss(side_data, _):
        return absorbDenseCenterSOSIntoSideProjected(direction, side_data, state_center_data, None, left, right,
                                                     state_center_data_conj)

    def sos(side_data, op_data):
        return absorbDenseCenterSOSIntoSideProjected(direction, side_data, state_center_data, op_data, left, right,
                                                     state_center_data_conj)
    return contractSparseTensors(_center_into_side_terms(direction, ss, sos), side, operator_center_data)


def formExpectationStage1(corner, side):
    """sparse.py:72-85 (handler arguments: corner tag, side tag)."""
    terms = {}
    f = formNormalizationStage1
    addStandardCompleteAndIdentityTerms(terms, f)
    terms.update({
        (TwoSiteOperator, Identity): lambda c, s: (c.matchesSideIdentityOnRightForStage1(), f),
        (Identity, TwoSiteOperator): lambda c, s: (s.matchesCornerIdentityOnLeftForStage1(), f),
        (TwoSiteOperator, TwoSiteOperator): lambda c, s: (c.matches(s), f),
        (TwoSiteOperatorCompressed, Identity): lambda c, s: (c.matchesSideIdentityOnRightForStage1(), f),
        (Identity, TwoSiteOperatorCompressed): lambda c, s: (s.matchesCornerIdentityOnLeftForStage1(), f),
        (TwoSiteOperatorCompressed, TwoSiteOperatorCompressed): lambda c, s: (c.matches(s), f),
    })
    return contractSparseTensors(terms, corner, side)


def formExpectationStage2(right, left):
    """sparse.py:86-99 (handler arguments: right tag, left tag)."""
    terms = {}
    f = formNormalizationStage2
    addStandardCompleteAndIdentityTerms(terms, f)
    terms.update({
        (TwoSiteOperator, Identity): lambda r, l: (r.matchesStage1IdentityOnLeft(), f),
        (Identity, TwoSiteOperator): lambda r, l: (l.matchesStage1IdentityOnRight(), f),
        (TwoSiteOperator, TwoSiteOperator): lambda r, l: (l.matches(r), f),
        (TwoSiteOperatorCompressed, Identity): lambda r, l: (r.matchesStage1IdentityOnLeft(), f),
        (Identity, TwoSiteOperatorCompressed): lambda r, l: (l.matchesStage1IdentityOnRight(), f),
        (TwoSiteOperatorCompressed, TwoSiteOperatorCompressed): lambda r, l: (l.matches(r), f),
    })
    return contractSparseTensors(terms, right, left)


STAGE3_RULES = {
    (Complete, Identity, Identity): lambda x, y, z: True,
    (Identity, Complete, Identity): lambda x, y, z: True,
    (Identity, Identity, OneSiteOperator): lambda x, y, z: True,
    (TwoSiteOperator, Identity, TwoSiteOperator): lambda x, y, z: x.matchesCenterForStage3(0, z),
    (Identity, TwoSiteOperator, TwoSiteOperator): lambda x, y, z: y.matchesCenterForStage3(1, z),
    (TwoSiteOperator, TwoSiteOperator, Identity): lambda x, y, z: x.matchesForStage3(y),
    (TwoSiteOperatorCompressed, TwoSiteOperatorCompressed, Identity): lambda x, y, z: x.matchesForStage3(y),
}


def stage3_terms(stage2_0, stage2_1, operator_center):
    """The compatible tag triples in the reference's enumeration order (sparse.py:116-127),
    as (stage2_0 data, stage2_1 data, operator data or None for an identity operator)."""
    terms = []
    for tags in itertools.product(stage2_0, stage2_1, operator_center):
        rule = STAGE3_RULES.get(tuple(type(t) for t in tags))
        if rule is None or not rule(*tags):
            continue
        x, y, z = tags
        terms.append((stage2_0[x], stage2_1[y], None if z == Identity() else operator_center[z]))
    return terms


def formExpectationStage3(stage2_0, stage2_1, operator_center):
    """sparse.py:100-161: (expectation multiplier, normalization multiplier)."""
    d, _, DataClass = getInformationFromOperatorCenter(operator_center)
    terms = stage3_terms(stage2_0, stage2_1, operator_center)
    normalization = stage3.Stage3Operator([(stage2_0[Identity()], stage2_1[Identity()], None)], physical_dimension=d)
    if terms:
        expectation = stage3.Stage3Operator(terms, physical_dimension=d).multiplier()
    else:
        from ...utils import Multiplier
        n = normalization.n

        def zero(v):
            return v.newZeros(v.shape)
        expectation = Multiplier((n, n), zero, 0, lambda: DataClass.newZeros((n, n)), 0)
    return expectation, normalization.multiplier()


def formExpectationAndNormalizationMultipliers(corners, sides, operator_center_data):
    """sparse.py:59-71."""
    return formExpectationStage3(
        formExpectationStage2(formExpectationStage1(corners[0], sides[0]), formExpectationStage1(corners[1], sides[1])),
        formExpectationStage2(formExpectationStage1(corners[2], sides[2]), formExpectationStage1(corners[3], sides[3])),
        operator_center_data)
