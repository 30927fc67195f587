"""Dense contractors of the 2D hot path.  Re-implements carcassonne/tensors/_2d/dense.py.

The index maps (which legs are joined, and the leg order of each result) are the
reference's and are given here as contraction plans for
carcassonne_amd.utils.formDataContractor; on HipData every pairwise step is one
cc_gemm_f64 call (plus cc_permute where the legs are not already in GEMM order).

Stage 3 -- the effective-Hamiltonian matvec -- is not run through a generated
contractor: formNormalizationStage3 / formDenseStage3 return Multipliers backed
by carcassonne_amd.stage3.Stage3Operator, which applies all terms of a solve in one
call to the HIP library.
"""
from ...data.cost_tracker import CostTracker, computeCostOfContracting
from ...utils import Join, Multiplier, formDataContractor, prepend, L, R, O
from ... import stage3

# corner C(0,1,2) <-> side T(3,4,5)  ->  (T0, T1, T2, (C3,T6), (C4,T7), C5)      dense.py:11-14
absorbDenseSideIntoCornerFromLeft = formDataContractor(
    [Join(0, [0, 1, 2], 1, [3, 4, 5])],
    [[(1, 0)], [(1, 1)], [(1, 2)], [(0, 3), (1, 6)], [(0, 4), (1, 7)], [(0, 5)]])

# corner C(3,4,5) <-> side T(0,1,2)  ->  ((C0,T6), (C1,T7), C2, T3, T4, T5)      dense.py:17-20
absorbDenseSideIntoCornerFromRight = formDataContractor(
    [Join(0, [3, 4, 5], 1, [0, 1, 2])],
    [[(0, 0), (1, 6)], [(0, 1), (1, 7)], [(0, 2)], [(1, 3)], [(1, 4)], [(1, 5)]])


def _center_groups(i):
    return [[(0, 0), (1, L(i))], [(0, 1), (2, L(i))], [(0, 2)], [(0, 3), (1, R(i))], [(0, 4), (2, R(i))],
            [(0, 5)], [(1, O(i))], [(2, O(i))]]


# side T6<->S_i, T7<->S*_i, S4<->S*4                                              dense.py:23-40
_ss = [formDataContractor([Join(0, 6, 1, i), Join(0, 7, 2, i), Join(1, 4, 2, 4)], _center_groups(i)) for i in range(4)]
# operator O0<->S*4, O1<->S4, then as SS                                          dense.py:52-70
_sos = [formDataContractor([Join(3, 0, 2, 4), Join(3, 1, 1, 4), Join(0, 6, 1, i), Join(0, 7, 2, i)], _center_groups(i))
        for i in range(4)]


def absorbDenseCenterSSIntoSide(direction, side, center, center_conj=None):
    if center_conj is None:
        center_conj = center.conj()
    return _ss[direction](side, center, center_conj)


def absorbDenseCenterSOSIntoSide(direction, side, state_center_data, operator_center_data, state_center_data_conj=None):
    if state_center_data_conj is None:
        state_center_data_conj = state_center_data.conj()
    return _sos[direction](side, state_center_data, state_center_data_conj, operator_center_data)


def absorbDenseCenterSOSIntoSideProjected(direction, side, state, operator, left, right, state_conj=None):
    """absorbDenseCenterSOSIntoSide (operator=None: absorbDenseCenterSSIntoSide) followed by the projection
    of both grown bonds onto fused truncation bases, without forming the grown side:

        out[k, k', o, o'] = sum left[(x, a, b), k] right[(y, c, e), k'] T[x, y, i, i'] v[.., s] O[s', s] v*[.., s']

    with the side's center legs (i, i') meeting the state's leg `direction`, (a, b) / (c, e) its legs L / R
    (ket, bra) and (o, o') its leg O.  The side's bond legs are in fused form (x, 1, 1) and (y, 1, 1); the
    result is (k, 1, 1, k', 1, 1, o, o').  Four GEMMs; the largest intermediate has x D^4 d k elements
    (the grown side would have x y D^6)."""
    i = direction
    sh = side.shape
    if sh[1] * sh[2] * sh[4] * sh[5] != 1:
        raise ValueError("the projected absorption needs fused side bonds (x, 1, 1), got " + str(tuple(sh)))
    if state_conj is None:
        state_conj = state.conj()
    x, y = sh[0], sh[3]
    T = side.join(0, 3, 6, 7)                                    # (x, y, i, i')
    v = state.join(i, L(i), R(i), O(i), 4)                       # (i, a, c, o, s)
    vb = state_conj.join(i, L(i), R(i), O(i), 4)                 # (i', b, e, o', s')
    if operator is not None:                                     # (O v)[.., s'] = sum_s O[s', s] v[.., s]
        v = operator.contractWith(v, (1,), (4,)).join(1, 2, 3, 4, 0)
    D_a, D_c = v.shape[1], v.shape[2]
    k, k2 = left.shape[1], right.shape[1]
    lt = left.split(x, D_a, D_a, k)                              # (x, a, b, k)
    rt = right.split(y, D_c, D_c, k2)                            # (y, c, e, k')
    X1 = lt.contractWith(T, (0,), (0,))                          # (a, b, k, y, i, i')
    X2 = X1.contractWith(v, (4, 0), (0, 1))                      # (b, k, y, i', c, o, s)
    X3 = X2.contractWith(vb, (3, 0, 6), (0, 1, 4))               # (k, y, c, o, e, o')
    out = X3.contractWith(rt, (1, 2, 4), (0, 1, 2))              # (k, o, o', k')
    o = out.shape[1]
    return out.join(0, 3, 1, 2).split(k, 1, 1, k2, 1, 1, o, o)


# corner C(3,4,5) <-> side T(0,1,2) -> ((C0,C1,C2), (T3,T4,T5), T6, T7)          dense.py:96-99
formNormalizationStage1 = formDataContractor(
    [Join(0, [3, 4, 5], 1, [0, 1, 2])],
    [[(0, 0), (0, 1), (0, 2)], [(1, 3), (1, 4), (1, 5)], [(1, 6)], [(1, 7)]])

# P0 <-> Q1 -> (Q0, P1, P2, Q2, P3, Q3)                                            dense.py:102-112
formNormalizationStage2 = formDataContractor(
    [Join(0, 0, 1, 1)],
    [[(1, 0)], [(0, 1)], [(0, 2)], [(1, 2)], [(0, 3)], [(1, 3)]])

# dense matrix of one term: s2_0(0,1) <-> s2_1(1,0), rows (s2_0 4,5; s2_1 4,5; O0),
# columns (s2_0 2,3; s2_1 2,3; O1)                                                dense.py:176-194
_form_matrix = formDataContractor(
    [Join(0, [0, 1], 1, [1, 0])],
    [[(0, 4), (0, 5), (1, 4), (1, 5), (2, 0)], [(0, 2), (0, 3), (1, 2), (1, 3), (2, 1)]])
_form_submatrix = formDataContractor(
    [Join(0, [0, 1], 1, [1, 0])],
    [[(0, 4), (0, 5), (1, 4), (1, 5)], [(0, 2), (0, 3), (1, 2), (1, 3)]])


def formDenseStage3_formMatrix_and_cost(stage2_0, stage2_1, site_operator):
    return (lambda: _form_matrix(stage2_0, stage2_1, site_operator),
            computeCostOfContracting(_form_matrix, stage2_0, stage2_1, site_operator))


def formDenseStage3_multiply_and_cost(stage2_0, stage2_1, site_operator):
    op = stage3.Stage3Operator([(stage2_0, stage2_1, site_operator)])
    return op.apply, op.cost_of_multiply


def formDenseStage3(stage2_0, stage2_1, site_operator):
    op = stage3.Stage3Operator([(stage2_0, stage2_1, site_operator)])
    return op.multiplier()


def formNormalizationStage3(stage2_0, stage2_1, center_identity):
    op = stage3.Stage3Operator([(stage2_0, stage2_1, None)], physical_dimension=center_identity.shape[0])
    return op.multiplier()


def formNormalizationMultiplier(corners, sides, center_identity):
    return formNormalizationStage3(
        formNormalizationStage2(formNormalizationStage1(corners[0], sides[0]), formNormalizationStage1(corners[1], sides[1])),
        formNormalizationStage2(formNormalizationStage1(corners[2], sides[2]), formNormalizationStage1(corners[3], sides[3])),
        center_identity)


def formNormalizationSubmatrix(corners, sides):
    return _form_submatrix(
        formNormalizationStage2(formNormalizationStage1(corners[0], sides[0]), formNormalizationStage1(corners[1], sides[1])),
        formNormalizationStage2(formNormalizationStage1(corners[2], sides[2]), formNormalizationStage1(corners[3], sides[3])))


__all__ = ["absorbDenseSideIntoCornerFromLeft", "absorbDenseSideIntoCornerFromRight", "absorbDenseCenterSSIntoSide",
           "absorbDenseCenterSOSIntoSide", "absorbDenseCenterSOSIntoSideProjected", "formDenseStage3", "formDenseStage3_formMatrix_and_cost",
           "formDenseStage3_multiply_and_cost", "formNormalizationMultiplier", "formNormalizationStage1",
           "formNormalizationStage2", "formNormalizationStage3", "formNormalizationSubmatrix"]
