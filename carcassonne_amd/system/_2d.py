"""The 2D system on the GPU.  Re-implements carcassonne/system/_2d.py (System, 19-569) with
device-resident data (HipData): constructors, consistency checks, state and operator
compression entry points, observables, growth (contractTowards), gauge moves, bandwidth
increase and the solve.

minimizeExpectation keeps the reference's relaxOver by default; solver="lanczos" runs the
residual-controlled FP64 Lanczos and solver="mixed" the FP16/BF16 precision form (both need
an instantiated fused shape, P = D0*D1 in {4, 9, 16, 25, 36}, d = 2).
"""
import copy as _copy

import numpy as np

from .base import BaseSystem
from ..compression import computeProductCompressor
from ..data import HipData
from ..errors import InvariantViolatedError
from ..sparse import (Identity, OneSiteOperator, TwoSiteOperator, TwoSiteOperatorCompressed, makeSimpleSparseOperator,
                      makeSparseOperator, mapOverSparseData, stripAllButIdentityFrom)
from ..tensors._2d.dense import formNormalizationMultiplier, formNormalizationSubmatrix
from ..tensors._2d.sparse import (absorbSparseCenterSOSIntoSide, absorbSparseSideIntoCornerFromLeft,
                                  absorbSparseSideIntoCornerFromRight, formExpectationAndNormalizationMultipliers)
from ..utils import L, O, R, Multiplier, computeCompressor, relaxOver


def sideFromCorner(corner_id, direction):
    """_2d.py:575-576."""
    return (corner_id + 1 - direction) % 4


def fusedTruncationBasis(corner, bond_legs, far_legs, new_dimension):
    """The (x D D) x k truncation basis of a grown fused corner bond, bond legs (x D, D, 1) with the old bond
    x and the absorbed side's center ket / bra legs D, from the corner's Identity channel matricized bond x
    far (cc_fused_truncation_f64); k = min(new_dimension, far)."""
    from .. import _lib
    from ..device import empty, ptr, stream
    sh = corner.shape
    b = [sh[j] for j in bond_legs]
    D = b[1]
    if b[2] != 1 or b[0] % D:
        raise ValueError("fused truncation needs a grown fused bond (x D, D, 1), got {}".format(tuple(b)))
    nx = b[0] // D
    far = int(np.prod([sh[j] for j in far_legs]))
    k = min(int(new_dimension), far)
    C = corner.join(tuple(bond_legs), tuple(far_legs)).tensor().contiguous()
    m = nx * D * D
    U, S = empty((m, k)), empty((far,))
    work = empty((_lib.query("cc_fused_truncation_work_size", nx, D, far),))
    _lib.call("cc_fused_truncation_f64", stream(), nx, D, far, k, ptr(C), ptr(U), ptr(S), ptr(work))
    return HipData(U)


def _project_corner(t, bond_legs, basis):
    """A corner channel with its fused bond legs (x D, D, 1) replaced by (k, 1, 1) through `basis`."""
    sh = t.shape
    k = basis.shape[1]
    if tuple(bond_legs) == (3, 4, 5):
        out = t.join((0, 1, 2), (3, 4, 5)).contractWith(basis, (1,), (0,), split_long_sums=True)
        return out.split(sh[0], sh[1], sh[2], k, 1, 1)
    out = basis.contractWith(t.join((0, 1, 2), (3, 4, 5)), (0,), (0,), split_long_sums=True)
    return out.split(k, 1, 1, sh[3], sh[4], sh[5])


class System(BaseSystem):
    # -- constructors ----------------------------------------------------------------
    @classmethod
    def newRandom(cls, makeOperator=None, DataClass=HipData, maximum_dimension=2, O=None, rng=None):
        """_2d.py:35-70 (random spoke sizes; ket/bra-symmetric sides and corners)."""
        assert not (makeOperator is not None and O is not None)
        rng = rng if rng is not None else np.random.default_rng()
        rdim = lambda: int(rng.integers(1, maximum_dimension + 1))
        spoke = (rdim(), rdim()) * 2
        sd = [rdim() for _ in range(4)]
        rnd = lambda *shape: DataClass.fromArray(rng.uniform(-1, 1, shape))
        sides = []
        for i in range(4):
            t = rnd(*((sd[i],) * 2 + (1,)) * 2 + (spoke[i],) * 2)
            t += t.join(1, 0, 2, 4, 3, 5, 7, 6)
            sides.append(t)
        corners = []
        for i in range(4):
            c = rnd(*(sides[L(i)].shape[3],) * 2 + (1,) + (sides[i].shape[0],) * 2 + (1,))
            c += c.join(1, 0, 2, 4, 3, 5)
            corners.append(c)
        d = O.shape[0] if O is not None else max(2, rdim())
        state = rnd(*spoke + (d,))
        if O is None:
            if makeOperator is None:
                O = rnd(d, d)
                O += O.join(1, 0)
            else:
                O = makeOperator(d)
        system = cls([{Identity(): c} for c in corners], [{Identity(): s} for s in sides], state,
                     {Identity(): DataClass.newIdentity(d), OneSiteOperator(None): O})
        system.assertDimensionsAreConsistent()
        system.assertNormalizationIsHermitian()
        system.assertHasNoNaNs()
        return system

    @classmethod
    def newTrivial(cls, operator_center_tensor, DataClass=HipData):
        """_2d.py:71-85."""
        d = None
        for data in operator_center_tensor.values():
            if data is not None:
                d = data.shape[0]
                break
        if d is None:
            raise ValueError("Operator tensor must have at least one non-identity component.")
        return cls([{Identity(): DataClass.newTrivial((1,) * 6)} for _ in range(4)],
                   [{Identity(): DataClass.newTrivial((1,) * 8)} for _ in range(4)],
                   DataClass.newFilled((1, 1, 1, 1, d), 1.0 / np.sqrt(d)), operator_center_tensor)

    @classmethod
    def newTrivialWithSimpleSparseOperator(cls, O=None, OO_UD=None, OO_LR=None):
        return cls.newTrivial(makeSimpleSparseOperator(O=O, OO_UD=OO_UD, OO_LR=OO_LR))

    @classmethod
    def newTrivialWithSparseOperator(cls, Os=(), OO_UDs=(), OO_LRs=()):
        return cls.newTrivial(makeSparseOperator(Os=Os, OO_UDs=OO_UDs, OO_LRs=OO_LRs))

    def __init__(self, corners, sides, state_center_data, operator_center_tensor, state_center_data_conj=None):
        BaseSystem.__init__(self)
        self.corners = list(corners)
        self.sides = list(sides)
        self.operator_center_tensor = operator_center_tensor
        self.state_center_data = state_center_data
        self.state_center_data_conj = state_center_data.conj() if state_center_data_conj is None else state_center_data_conj
        self.just_increased_bandwidth = False
        self.solver = "relaxover"
        self.solver_options = {}
        self.last_solve = {}
        # start of the state compressor's iteration: "random" (the reference's, compression.py:35) or "svd"
        self.compressor_start = "random"

    def __copy__(self):
        s = System(list(self.corners), list(self.sides), self.state_center_data, dict(self.operator_center_tensor),
                   self.state_center_data_conj)
        s.solver, s.solver_options = self.solver, dict(self.solver_options)
        s.compressor_start = self.compressor_start
        return s

    # -- checks (_2d.py:132-190) ---------------------------------------------------------
    def assertDimensionsAreConsistent(self):
        S = self.state_center_data
        assert S.shape == self.state_center_data_conj.shape
        if S.shape[0] != S.shape[2]:
            raise AssertionError("state center's left and right dimensions do not agree ({} != {})".format(S.shape[0], S.shape[2]))
        if S.shape[1] != S.shape[3]:
            raise AssertionError("state center's up and down dimensions do not agree ({} != {})".format(S.shape[1], S.shape[3]))
        for i, side in enumerate(self.sides):
            n = side[Identity()]
            if n.ndim != 8:
                raise AssertionError("for side {} the normalization data has rank {} instead of rank 8".format(i, n.ndim))
            for tag, data in side.items():
                if data.shape != n.shape:
                    raise AssertionError("for side {} the data tagged with {} does not match the shape of the data "
                                         "tagged with Identity() ({} != {})".format(i, tag, data.shape, n.shape))
            sh = n.shape
            for d in (0, 3, 6):
                if sh[d] != sh[d + 1]:
                    raise AssertionError("side {}'s dimension {} does not match its dimension {}".format(i, d, d + 1))
            if sh[0] != sh[3]:
                raise AssertionError("side {}'s left and right dimensions do not agree".format(i))
            if sh[6] != S.shape[i]:
                raise AssertionError("side {}'s center-facing dimensions do not match the corresponding state "
                                     "dimension ({} != {})".format(i, sh[6], S.shape[i]))
        for i, corner in enumerate(self.corners):
            n = corner[Identity()]
            if n.ndim != 6:
                raise AssertionError("for corner {} the normalization data has rank {} instead of rank 6".format(i, n.ndim))
            for tag, data in corner.items():
                if data.shape != n.shape:
                    raise AssertionError("for corner {} the data tagged with {} does not match the shape of the data "
                                         "tagged with Identity()".format(i, tag))
            sh = n.shape
            for d in (0, 3):
                if sh[d] != sh[d + 1]:
                    raise AssertionError("corner {}'s dimension {} does not match its dimension {}".format(i, d, d + 1))
            if sh[3] != self.sides[i][Identity()].shape[0]:
                raise AssertionError("corner {}'s right dimensions do not match side {}'s left dimensions".format(i, i))
            if sh[0] != self.sides[L(i)][Identity()].shape[3]:
                raise AssertionError("corner {}'s left dimensions do not match side {}'s right dimensions".format(i, L(i)))

    def assertHasNoNaNs(self):
        for i, corner in enumerate(self.corners):
            for tag, data in corner.items():
                if data.hasNaN():
                    raise AssertionError("corner {} has a NaN in component {}".format(i, tag))
        for i, side in enumerate(self.sides):
            for tag, data in side.items():
                if data.hasNaN():
                    raise AssertionError("side {} has a NaN in component {}".format(i, tag))
        if self.state_center_data.hasNaN():
            raise AssertionError("state center has a NaN")
        for tag, data in self.operator_center_tensor.items():
            if data is not None and data.hasNaN():
                raise AssertionError("operator center has a NaN in component {}".format(tag))

    def assertNormalizationIsHermitian(self):
        for i in range(4):
            s = self.sides[i][Identity()]
            if not s.allcloseTo(s.join(1, 0, 2, 4, 3, 5, 7, 6)):
                raise AssertionError("side {} is not hermitian".format(i))
            c = self.corners[i][Identity()]
            if not c.allcloseTo(c.join(1, 0, 2, 4, 3, 5)):
                raise AssertionError("corner {} is not hermitian".format(i))

    # -- compression of the corner-side bonds (_2d.py:191-228) -----------------------------
    def compressCornerStateTowards(self, corner_id, direction, new_dimension):
        if direction == 0:
            return self.compressCornerStateTowardsLeft(corner_id, new_dimension)
        elif direction == 1:
            return self.compressCornerStateTowardsRight(corner_id, new_dimension)
        raise ValueError("compression direction must be 0 or 1, not " + str(direction))

    def compressCornerStateTowardsLeft(self, corner_id, new_dimension):
        side_joined = self.sides[L(corner_id)][Identity()].join((0, 1, 2, 6, 7), 3, 4, 5)
        corner_joined = self.corners[corner_id][Identity()].join(0, 1, 2, (3, 4, 5))
        U = computeProductCompressor(side_joined, corner_joined, new_dimension, start=self.compressor_start)
        self.sides[L(corner_id)] = mapOverSparseData(lambda t: t.absorbMatrixAt(3, U).absorbMatrixAt(4, U.conj()),
                                                     self.sides[L(corner_id)])
        self.corners[corner_id] = mapOverSparseData(lambda t: t.absorbMatrixAt(0, U.conj()).absorbMatrixAt(1, U),
                                                    self.corners[corner_id])

    def compressCornerStateTowardsRight(self, corner_id, new_dimension):
        corner_joined = self.corners[corner_id][Identity()].join((0, 1, 2), 3, 4, 5)
        side_joined = self.sides[corner_id][Identity()].join(0, 1, 2, (3, 4, 5, 6, 7))
        U = computeProductCompressor(corner_joined, side_joined, new_dimension, start=self.compressor_start)
        self.corners[corner_id] = mapOverSparseData(lambda t: t.absorbMatrixAt(3, U).absorbMatrixAt(4, U.conj()),
                                                    self.corners[corner_id])
        self.sides[corner_id] = mapOverSparseData(lambda t: t.absorbMatrixAt(0, U.conj()).absorbMatrixAt(1, U),
                                                  self.sides[corner_id])

    def compressCornerTwoSiteOperatorTowards(self, corner_id, direction, new_dimension, normalize=False):
        """_2d.py:229-363: fold the corner's open two-site channels of `direction` and its compressed
        channel into one TwoSiteOperatorCompressed(direction) channel of new_dimension along the
        operator leg (axis 3*direction + 2), and the matching channels of the side it meets into
        TwoSiteOperatorCompressed(1 - direction).  The Gram matrix is block diagonal (open channels |
        compressed channel) as in the reference; the open channels are reshaped to their own count (the
        reference reshapes to the total old dimension, which only works before any compression)."""
        axis = 3 * direction + 2
        new_corner, sparse, pos_map, old_comp = {}, [], {}, None
        for tag, data in self.corners[corner_id].items():
            if isinstance(tag, TwoSiteOperator) and tag.direction == direction:
                assert tag.position not in pos_map
                pos_map[tag.position] = len(sparse)
                sparse.append(data)
            elif isinstance(tag, TwoSiteOperatorCompressed) and tag.direction == direction:
                assert old_comp is None
                old_comp = data
            else:
                new_corner[tag] = data
        n1 = len(sparse)
        old_dimension = n1 + (old_comp.shape[axis] if old_comp is not None else 0)
        if old_dimension == 0:
            return
        blocks = []
        if n1:
            collected = type(sparse[0]).newCollected(sparse).dropUnitAxis(axis + 1)  # (n1, others)
            flat = collected.join(0, list(range(1, collected.ndim)))
            blocks.append(flat.conj().contractWith(flat, (1,), (1,)))
        if old_comp is not None:
            f = old_comp.fold(axis)
            blocks.append(f.conj().contractWith(f, (1,), (1,)))
        M = blocks[0] if len(blocks) == 1 else blocks[0].directSumWith(blocks[1])
        corner_mult, side_mult_conj = computeCompressor(
            old_dimension, new_dimension,
            Multiplier((old_dimension,) * 2, lambda v: M.contractWith(v, (1,), (0,)), old_dimension ** 2,
                       lambda: M, 0),
            None, normalize)
        side_mult = side_mult_conj.conj()

        def compress(part, mult, comp_old, ax):
            out = None
            if part is not None:
                t = mult[:, :n1].contractWith(part, (1,), (0,))                       # (new, others)
                perm = list(range(1, t.ndim))
                perm.insert(ax, 0)
                out = t.transpose(perm)
            if comp_old is not None:
                t2 = comp_old.absorbMatrixAt(ax, mult[:, n1:])
                out = t2 if out is None else out + t2
            return out

        new_corner[TwoSiteOperatorCompressed(direction)] = compress(collected if n1 else None, corner_mult,
                                                                    old_comp, axis)
        self.corners[corner_id] = new_corner
        side_id = sideFromCorner(corner_id, direction)
        sdir = 1 - direction
        saxis = 3 * sdir + 2
        parts, old_side, new_side = [None] * n1, None, {}
        for tag, data in self.sides[side_id].items():
            if isinstance(tag, TwoSiteOperator) and tag.direction == sdir:
                parts[pos_map[tag.position]] = data
            elif isinstance(tag, TwoSiteOperatorCompressed) and tag.direction == sdir:
                assert old_comp is not None and old_side is None
                old_side = data
            else:
                new_side[tag] = data
        assert all(p is not None for p in parts)
        spart = type(parts[0]).newCollected(parts).dropUnitAxis(saxis + 1) if n1 else None
        new_side[TwoSiteOperatorCompressed(sdir)] = compress(spart, side_mult, old_side, saxis)
        self.sides[side_id] = new_side

    # -- observables (_2d.py:364-431) ---------------------------------------------------------
    def computeScalarUsingMultiplier(self, multiply):
        return self.state_center_data_conj.contractWith(multiply(self.state_center_data), range(5), range(5)).extractScalar()

    def computeExpectationAndNormalization(self, operator_center_tensor=None):
        multiplyExpectation, multiplyNormalization = self.formExpectationAndNormalizationMultipliers(operator_center_tensor)
        unnormalized = self.computeScalarUsingMultiplier(multiplyExpectation)
        normalization = self.computeScalarUsingMultiplier(multiplyNormalization)
        return unnormalized / normalization, normalization

    def computeExpectation(self, operator_center_tensor=None):
        return self.computeExpectationAndNormalization(operator_center_tensor)[0]

    def computeExpectationWithoutCenter(self):
        return self.computeExpectationAndNormalizationWithoutCenter()[0]

    def computeExpectationAndNormalizationWithoutCenter(self):
        return self.computeExpectationAndNormalization({
            tag: value for (tag, value) in self.operator_center_tensor.items()
            if tag is Identity() or (isinstance(tag, TwoSiteOperator) and tag.position == 0 and tag.direction in (2, 3))})

    def computeCenterSiteExpectation(self):
        return self.computeExpectation() - self.computeExpectationWithoutCenter()

    def computeNormalization(self):
        return self.computeScalarUsingMultiplier(self.formNormalizationMultiplier())

    def computeUnnormalizedExpectation(self):
        return self.computeScalarUsingMultiplier(self.formExpectationMultiplier())

    def computeOneSiteExpectation(self):
        """_2d.py:396-428."""
        expectation = 0
        stripped = self.stripExpectationEnvironment()
        for tag, value in self.operator_center_tensor.items():
            if isinstance(tag, OneSiteOperator):
                system = _copy.copy(stripped)
                system.operator_center_tensor[tag] = value
                expectation += system.computeExpectation()
            if isinstance(tag, TwoSiteOperator):
                if tag.position != 0:
                    continue
                for first, second in ((0, 2), (1, 3)):
                    if tag.direction == first:
                        other = tag.withNewDirectionAndPosition(second, 0)
                        system = _copy.copy(stripped)
                        system.operator_center_tensor[tag] = value
                        system.operator_center_tensor[other] = self.operator_center_tensor[other]
                        system.contractTowards(first)
                        expectation += system.computeExpectation()
        return expectation

    def stripExpectationEnvironment(self):
        s = type(self)([stripAllButIdentityFrom(c) for c in self.corners], [stripAllButIdentityFrom(x) for x in self.sides],
                       self.state_center_data, stripAllButIdentityFrom(self.operator_center_tensor),
                       self.state_center_data_conj)
        s.solver, s.solver_options = self.solver, dict(self.solver_options)
        return s

    # -- growth (_2d.py:432-454) -----------------------------------------------------------
    def contractTowards(self, direction):
        tensor_to_contract, _, matrix_to_absorb = self.state_center_data.normalizeAxis(O(direction))
        self.contractUnnormalizedTowards(direction, tensor_to_contract)
        self.setStateCenter(self.state_center_data.normalizeAxis(direction)[0].absorbMatrixAt(direction, matrix_to_absorb))

    def contractNormalizedTowards(self, direction, state_center_data):
        self.contractUnnormalizedTowards(direction, state_center_data.normalizeAxis(O(direction))[0])

    def contractUnnormalizedTowards(self, direction, state_center_data=None, state_center_data_conj=None):
        if state_center_data is None:
            state_center_data = self.state_center_data
            state_center_data_conj = self.state_center_data_conj
        if state_center_data_conj is None:
            state_center_data_conj = state_center_data.conj()
        self.corners[direction] = absorbSparseSideIntoCornerFromLeft(self.corners[direction], self.sides[L(direction)])
        self.sides[direction] = absorbSparseCenterSOSIntoSide(direction, self.sides[direction], state_center_data,
                                                              self.operator_center_tensor, state_center_data_conj)
        self.corners[R(direction)] = absorbSparseSideIntoCornerFromRight(self.corners[R(direction)], self.sides[R(direction)])
        if self.just_increased_bandwidth:
            raise InvariantViolatedError("Contracting the current center would blow up the condition number of the "
                                         "normalization matrix;  optimize it or replace it first.")

    # -- the sweep step on fused environment bonds ------------------------------------------------
    def contractAndCompressTowards(self, direction, new_dimension):
        """contractTowards(direction) followed by the compression of the two bonds it grows, for an
        environment whose bonds are fused ((chi, 1, 1): ket, bra and operator legs in one index, the layout
        of the synthetic environments): corners `direction` and R(direction) absorb their sides as in
        contractUnnormalizedTowards (system/_2d.py:443-454); each grown corner bond (chi, D, D) is cut to a
        truncation basis (cc_fused_truncation_f64: the leading ket<->bra-symmetric left singular vectors of
        the corner's Identity channel, as compressCornerStateTowards takes its compressor from the Identity
        channel, system/_2d.py:191-228); every channel of both corners is projected onto it; and side
        `direction` absorbs the center already projected onto both bases
        (absorbSparseCenterSOSIntoSideProjected), so the grown side, (chi D^2)^2 D^2 elements per channel,
        never exists.  Bonds come back as (k, 1, 1) with k = min(new_dimension, chi).  With k equal to the
        corner's rank bound chi the Identity-channel products of both bonds are kept exactly."""
        from ..tensors._2d.sparse import absorbSparseCenterSOSIntoSideProjected
        if self.just_increased_bandwidth:
            raise InvariantViolatedError("Contracting the current center would blow up the condition number of the "
                                         "normalization matrix;  optimize it or replace it first.")
        import torch
        from ..device import fork, join
        i = direction
        iso, _, den = self.state_center_data.normalizeAxis(O(i))
        # the two corners and their truncation bases are independent: the right one on a side stream
        # (the bases' 64 x 64 eigen-decompositions run on one workgroup each)
        side, produced = fork(), []
        try:
            with torch.cuda.stream(side):
                corner_r = absorbSparseSideIntoCornerFromRight(self.corners[R(i)], self.sides[R(i)])
                right = fusedTruncationBasis(corner_r[Identity()], (0, 1, 2), (3, 4, 5), new_dimension)
                produced = list(corner_r.values()) + [right]
            corner_l = absorbSparseSideIntoCornerFromLeft(self.corners[i], self.sides[L(i)])
            left = fusedTruncationBasis(corner_l[Identity()], (3, 4, 5), (0, 1, 2), new_dimension)
        finally:
            join(side, produced)
        self.sides[i] = absorbSparseCenterSOSIntoSideProjected(i, self.sides[i], iso, self.operator_center_tensor,
                                                               left, right, iso.conj())
        self.corners[i] = mapOverSparseData(lambda t: _project_corner(t, (3, 4, 5), left), corner_l)
        self.corners[R(i)] = mapOverSparseData(lambda t: _project_corner(t, (0, 1, 2), right), corner_r)
        self.setStateCenter(self.state_center_data.normalizeAxis(i)[0].absorbMatrixAt(i, den))

    # -- multipliers (_2d.py:455-488) --------------------------------------------------------
    def formExpectationAndNormalizationMultipliers(self, operator_center_tensor=None):
        if operator_center_tensor is None:
            operator_center_tensor = self.operator_center_tensor
        return formExpectationAndNormalizationMultipliers(self.corners, self.sides, operator_center_tensor)

    def formExpectationMultiplier(self):
        return self.formExpectationAndNormalizationMultipliers()[0]

    def formExpectationMatrix(self):
        return self.formExpectationMultiplier().formMatrix()

    def formNormalizationMultiplier(self):
        return formNormalizationMultiplier(tuple(c[Identity()] for c in self.corners), tuple(s[Identity()] for s in self.sides),
                                           self.operator_center_tensor[Identity()])

    def formNormalizationMatrix(self):
        return self.formNormalizationMultiplier().formMatrix()

    def formNormalizationSubmatrix(self):
        return formNormalizationSubmatrix(tuple(c[Identity()] for c in self.corners), tuple(s[Identity()] for s in self.sides))

    def increaseBandwidth(self, direction, by=None, to=None, do_as_much_as_possible=False, enlargeners=None):
        if direction not in (0, 1):
            raise ValueError("Direction for bandwidth increase must be either 0 (for horizontal axes) or 1 (for "
                             "vertical axes), not {}.".format(direction))
        return self._increaseBandwidth(direction, by, to, do_as_much_as_possible, enlargeners)

    def increaseBandwidthAndThenNormalize(self, direction, by=None, to=None):
        self.increaseBandwidth(direction, by, to)
        self.normalize()

    # -- the solve (_2d.py:489-502) ------------------------------------------------------------
    def minimizeExpectation(self):
        H, N = self.formExpectationAndNormalizationMultipliers()
        if self.solver == "relaxover":
            stats = {}
            result = relaxOver(self.state_center_data, H, N, maximum_number_of_multiplications=100, stats=stats)
            self.last_solve = stats
            self.setStateCenter(result)
            return
        from ..solver import GroundStateSolver
        opts = dict(self.solver_options)
        if self.solver == "lanczos":
            solver = GroundStateSolver(H.operator, N.operator)
            x, theta, resid, iters = solver.solve(self.state_center_data, **opts)
            self.last_solve = dict(energy=theta, residual=resid, iterations=iters)
        elif self.solver == "mixed":
            from ..precision import MixedSolver
            dtype = opts.pop("dtype", "f16")
            solver = MixedSolver(H.operator, N.operator, dtype=dtype)
            x, rho, info = solver.solve(self.state_center_data, **opts)
            self.last_solve = dict(info, energy=rho)
        else:
            raise ValueError("unknown solver " + str(self.solver))
        self.setStateCenter(HipData(x.reshape(self.state_center_data.shape)))

    def minimizeExpectationUsingFullEigensolver(self):
        """The exact lowest generalized eigenpair: residual-controlled Lanczos run to 1e-12 where the
        operator has a fused kernel, else a dense generalized eigensolve of the formed H and N."""
        from ..solver import GroundStateSolver, dense_lowest_eigenpair
        H, N = self.formExpectationAndNormalizationMultipliers()
        op = getattr(H, "operator", None)     # absent when no stage-3 term survives (H = 0)
        if op is None or not op.use_fused:
            x, _ = dense_lowest_eigenpair(H.formMatrix(), N.formMatrix())
            self.setStateCenter(HipData(x.reshape(self.state_center_data.shape)))
            return
        x, theta, resid, iters = GroundStateSolver(op, N.operator).solve(
            self.state_center_data, tol_residual=1e-12, tol_energy=0.0, max_iter=op.n)
        self.setStateCenter(HipData(x.reshape(self.state_center_data.shape)))

    # -- gauge moves (_2d.py:503-549) ------------------------------------------------------------
    def normalize(self):
        for corner_id in range(4):
            for direction in range(2):
                self.normalizeCornerAndDenormalizeSide(corner_id, direction)
        for side_id in range(4):
            self.normalizeSideAndDenormalizeCenter(side_id)

    def normalizeCenterAndDenormalizeSide(self, direction):
        normalizer, denormalizer1 = self.state_center_data.normalizeAxis(direction, True)
        self.state_center_data = self.state_center_data.absorbMatrixAt(direction, normalizer)
        self.state_center_data_conj = self.state_center_data.conj()
        denormalizer2 = denormalizer1.conj()
        self.sides[direction] = {tag: d.absorbMatrixAt(6, denormalizer1).absorbMatrixAt(7, denormalizer2)
                                 for tag, d in self.sides[direction].items()}

    def normalizeCornerAndDenormalizeSide(self, corner_id, direction):
        side_id = sideFromCorner(corner_id, direction)
        ca1, ca2 = direction * 3, direction * 3 + 1
        sa1, sa2 = (1 - direction) * 3, (1 - direction) * 3 + 1
        n1, d1 = self.corners[corner_id][Identity()].normalizeAxis(ca1, True)
        n2, d2 = n1.conj(), d1.conj()
        self.corners[corner_id] = {tag: c.absorbMatrixAt(ca1, n1).absorbMatrixAt(ca2, n2)
                                   for tag, c in self.corners[corner_id].items()}
        self.sides[side_id] = {tag: s.absorbMatrixAt(sa1, d1).absorbMatrixAt(sa2, d2)
                               for tag, s in self.sides[side_id].items()}

    def normalizeSideAndDenormalizeCenter(self, side_id):
        n1, dc = self.sides[side_id][Identity()].normalizeAxis(6, True)
        n2 = n1.conj()
        self.sides[side_id] = {tag: s.absorbMatrixAt(6, n1).absorbMatrixAt(7, n2) for tag, s in self.sides[side_id].items()}
        self.state_center_data = self.state_center_data.absorbMatrixAt(side_id, dc)
        self.state_center_data_conj = self.state_center_data.conj()

    def setStateCenter(self, state_center_data, state_center_data_conj=None):
        self.state_center_data = state_center_data
        self.state_center_data_conj = state_center_data.conj() if state_center_data_conj is None else state_center_data_conj
        self.just_increased_bandwidth = False


__all__ = ["System", "sideFromCorner"]
