"""The 2D system, its observables and the sweep driver (oracle).

Restates carcassonne/system/_2d.py (System: 19-569), carcassonne/system/base.py
(sweep/run driver 65-111 and _increaseBandwidth 114-160) and
carcassonne/policies.py (80-242).  Randomness is drawn from an explicit
numpy Generator (the reference uses unseeded global RNGs).

Pinned by tests/test_oracle_system.py: identity-operator walks (E =
width*height exactly, reference test_system.py:308-358), Z-field product-state
walks (test_system.py:360-439), matrix-free == dense (test_system.py:222-275),
gauge invariance of normalize* (test_system.py:276-305), compression exactness
at full rank (test_system.py:40-81), and the 2D-in-1D known answers
(-1, -1.0000250001562545, 1/4 - ln 2; test_simulator_2d_in_1d.py:21-60).
"""
import copy as _copy

import numpy as np
from scipy.linalg import eigh

from . import dense, linalg, solver, sparse2d
from .tags import (Identity, OneSiteOperator, TwoSiteOperator, TwoSiteOperatorCompressed, L, R, O,
                   make_simple_sparse_operator, make_sparse_operator, strip_all_but_identity)


class InvariantViolatedError(Exception):
    pass


def side_from_corner(corner_id, direction):
    """sideFromCorner (_2d.py:575-576)."""
    return (corner_id + 1 - direction) % 4


def _map(f, sparse):
    return {t: f(d) for t, d in sparse.items()}


class System:
    POLICY_NAMES = ["bandwidth increase", "contraction", "operator compression", "run convergence",
                    "post-contraction hook", "pre-optimization hook", "post-optimization hook",
                    "state compression", "sweep convergence"]

    def __init__(self, corners, sides, state, op_center, state_conj=None, rng=None):
        self.corners = list(corners)
        self.sides = list(sides)
        self.op_center = op_center
        self.state = state
        self.state_conj = state.conj() if state_conj is None else state_conj
        self.just_increased_bandwidth = False
        self.rng = rng if rng is not None else np.random.default_rng(0)
        self.policies = {n: None for n in self.POLICY_NAMES}
        self.solver = "relaxover"   # or "eigh"
        self.last_solve = {}
        self.compressor_start = "random"   # or "svd" (linalg.compressor_start)

    # -- constructors ---------------------------------------------------------------
    @classmethod
    def newTrivial(cls, op_center, rng=None):
        """_2d.py:71-85 (dtype follows the operator)."""
        d = next(o for o in op_center.values() if o is not None).shape[0]
        dt = np.result_type(*[o.dtype for o in op_center.values()], np.float64)
        return cls([{Identity(): np.ones((1,) * 6, dtype=dt)} for _ in range(4)],
                   [{Identity(): np.ones((1,) * 8, dtype=dt)} for _ in range(4)],
                   np.full((1, 1, 1, 1, d), 1.0 / np.sqrt(d), dtype=dt), op_center, rng=rng)

    @classmethod
    def newTrivialWithSimpleSparseOperator(cls, O=None, OO_UD=None, OO_LR=None, rng=None):
        return cls.newTrivial(make_simple_sparse_operator(O=O, OO_UD=OO_UD, OO_LR=OO_LR), rng=rng)

    @classmethod
    def newTrivialWithSparseOperator(cls, Os=(), OO_UDs=(), OO_LRs=(), rng=None):
        return cls.newTrivial(make_sparse_operator(Os=Os, OO_UDs=OO_UDs, OO_LRs=OO_LRs), rng=rng)

    @classmethod
    def newRandom(cls, rng, maximum_dimension=2, O=None, makeOperator=None, complex_=False):
        """_2d.py:35-70: random spoke sizes, Hermitian (ket<->bra) sides and corners."""
        rnd = (lambda *s: rng.uniform(-1, 1, s) + 1j * rng.uniform(-1, 1, s)) if complex_ else \
              (lambda *s: rng.uniform(-1, 1, s))
        rdim = lambda: int(rng.integers(1, maximum_dimension + 1))
        spoke = (rdim(), rdim()) * 2
        sd = [rdim() for _ in range(4)]
        sides = []
        for i in range(4):
            t = rnd(sd[i], sd[i], 1, sd[i], sd[i], 1, spoke[i], spoke[i])
            t = t + t.transpose(1, 0, 2, 4, 3, 5, 7, 6).conj()
            sides.append(t)
        corners = []
        for i in range(4):
            c = rnd(sides[L(i)].shape[3], sides[L(i)].shape[3], 1, sides[i].shape[0], sides[i].shape[0], 1)
            c = c + c.transpose(1, 0, 2, 4, 3, 5).conj()
            corners.append(c)
        d = O.shape[0] if O is not None else max(2, rdim())
        state = rnd(*spoke, d)
        if O is None:
            if makeOperator is None:
                O = rnd(d, d)
                O = O + O.T.conj()
            else:
                O = makeOperator(d)
        op = {Identity(): np.identity(d), OneSiteOperator(None): O}
        return cls([{Identity(): c} for c in corners], [{Identity(): s} for s in sides], state, op, rng=rng)

    def __copy__(self):
        s = System(list(self.corners), list(self.sides), self.state, dict(self.op_center),
                   self.state_conj, rng=self.rng)
        s.solver = self.solver
        return s

    # -- consistency checks (_2d.py:132-190) ----------------------------------------
    def assertNormalizationIsHermitian(self):
        for i in range(4):
            s = self.sides[i][Identity()]
            if not np.allclose(s, s.transpose(1, 0, 2, 4, 3, 5, 7, 6).conj()):
                raise AssertionError("side {} is not hermitian".format(i))
            c = self.corners[i][Identity()]
            if not np.allclose(c, c.transpose(1, 0, 2, 4, 3, 5).conj()):
                raise AssertionError("corner {} is not hermitian".format(i))

    # -- state --------------------------------------------------------------------
    def setStateCenter(self, state, state_conj=None):
        self.state = state
        self.state_conj = state.conj() if state_conj is None else state_conj
        self.just_increased_bandwidth = False

    # -- growth (_2d.py:435-454) ------------------------------------------------------
    def contractTowards(self, direction):
        iso, _, den = linalg.normalize_axis(self.state, O(direction))
        self.contractUnnormalizedTowards(direction, iso)
        self.setStateCenter(linalg.absorb_matrix_at(linalg.normalize_axis(self.state, direction)[0], direction, den))

    def contractUnnormalizedTowards(self, direction, S=None, Sc=None):
        if S is None:
            S, Sc = self.state, self.state_conj
        if Sc is None:
            Sc = S.conj()
        self.corners[direction] = sparse2d.absorb_side_into_corner_from_left(self.corners[direction], self.sides[L(direction)])
        self.sides[direction] = sparse2d.absorb_center_sos_into_side(direction, self.sides[direction], S, self.op_center, Sc)
        self.corners[R(direction)] = sparse2d.absorb_side_into_corner_from_right(self.corners[R(direction)], self.sides[R(direction)])
        if self.just_increased_bandwidth:
            raise InvariantViolatedError("Contracting the current center would blow up the condition number "
                                         "of the normalization matrix;  optimize it or replace it first.")

    # -- compression (_2d.py:191-228) ---------------------------------------------------
    def compressCornerStateTowards(self, corner_id, direction, new_dimension):
        if direction == 0:
            return self.compressCornerStateTowardsLeft(corner_id, new_dimension)
        if direction == 1:
            return self.compressCornerStateTowardsRight(corner_id, new_dimension)
        raise ValueError("compression direction must be 0 or 1, not " + str(direction))

    def compressCornerStateTowardsLeft(self, c, new_dimension):
        side = linalg.join(self.sides[L(c)][Identity()], (0, 1, 2, 6, 7), 3, 4, 5)
        corner = linalg.join(self.corners[c][Identity()], 0, 1, 2, (3, 4, 5))
        U = linalg.product_compressor(side, corner, new_dimension, self.rng, start=self.compressor_start)
        self.sides[L(c)] = _map(lambda t: linalg.absorb_matrix_at(linalg.absorb_matrix_at(t, 3, U), 4, U.conj()), self.sides[L(c)])
        self.corners[c] = _map(lambda t: linalg.absorb_matrix_at(linalg.absorb_matrix_at(t, 0, U.conj()), 1, U), self.corners[c])

    def compressCornerStateTowardsRight(self, c, new_dimension):
        corner = linalg.join(self.corners[c][Identity()], (0, 1, 2), 3, 4, 5)
        side = linalg.join(self.sides[c][Identity()], 0, 1, 2, (3, 4, 5, 6, 7))
        U = linalg.product_compressor(corner, side, new_dimension, self.rng, start=self.compressor_start)
        self.corners[c] = _map(lambda t: linalg.absorb_matrix_at(linalg.absorb_matrix_at(t, 3, U), 4, U.conj()), self.corners[c])
        self.sides[c] = _map(lambda t: linalg.absorb_matrix_at(linalg.absorb_matrix_at(t, 0, U.conj()), 1, U), self.sides[c])

    def compressCornerTwoSiteOperatorTowards(self, corner_id, direction, new_dimension, normalize=False):
        """_2d.py:229-363: compress the corner's open two-site channels of `direction` (and its existing
        compressed channel) into one TwoSiteOperatorCompressed(direction) channel of new_dimension
        along the operator leg, with the matching side channels.  The Gram matrix is block diagonal
        (open channels | compressed channel), as in the reference.  The open channels are reshaped to
        their own count (the reference reshapes them to the total old dimension, which only works when
        no compressed channel exists yet)."""
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
        collected = np.array([d.ravel() for d in sparse]).reshape(n1, -1)
        m1 = collected.conj() @ collected.T
        if old_comp is not None:
            f = linalg.join(old_comp, axis, [i for i in range(old_comp.ndim) if i != axis])
            m2 = f.conj() @ f.T
        else:
            m2 = np.zeros((0, 0))
        M = np.zeros((old_dimension, old_dimension), dtype=np.result_type(m1, m2))
        M[:n1, :n1] = m1
        M[n1:, n1:] = m2
        corner_mult, side_mult_conj = linalg.compute_compressor(old_dimension, new_dimension, M, normalize)
        side_mult = side_mult_conj.conj()
        shape = list(self.corners[corner_id][Identity()].shape)
        del shape[axis]

        def compress(part, mult, comp_old, ax):
            out = None
            if part is not None:
                t = np.tensordot(mult[:, :n1], part, (1, 0))          # (new, others)
                out = np.moveaxis(t, 0, ax)
            if comp_old is not None:
                t2 = linalg.absorb_matrix_at(comp_old, ax, mult[:, n1:])
                out = t2 if out is None else out + t2
            return out

        part = collected.reshape([n1] + shape) if n1 else None
        new_corner[TwoSiteOperatorCompressed(direction)] = compress(part, corner_mult, old_comp, axis)
        self.corners[corner_id] = new_corner
        side_id = side_from_corner(corner_id, direction)
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
        spart = np.array([np.squeeze(d, axis=saxis) for d in parts]) if n1 else None
        new_side[TwoSiteOperatorCompressed(sdir)] = compress(spart, side_mult, old_side, saxis)
        self.sides[side_id] = new_side

    # -- operators and observables (_2d.py:367-431,455-459) -----------------------------
    def formOperator(self, op_center=None):
        return sparse2d.form_multipliers(self.corners, self.sides, self.op_center if op_center is None else op_center)

    def _scalar(self, y):
        """computeScalarUsingMultiplier (_2d.py:429-431): sum over all legs of S* (M S)."""
        return np.sum(self.state_conj * y)

    def computeExpectationAndNormalization(self, op_center=None):
        op = self.formOperator(op_center)
        e = self._scalar(op.apply_H(self.state))
        n = self._scalar(op.apply_N(self.state))
        return e / n, n

    def computeExpectation(self, op_center=None):
        return self.computeExpectationAndNormalization(op_center)[0]

    def computeNormalization(self):
        return self.computeExpectationAndNormalization()[1]

    def computeExpectationWithoutCenter(self):
        return self.computeExpectation({t: v for t, v in self.op_center.items()
                                        if t is Identity() or (isinstance(t, TwoSiteOperator) and t.position == 0 and t.direction in (2, 3))})

    def computeCenterSiteExpectation(self):
        return self.computeExpectation() - self.computeExpectationWithoutCenter()

    def stripExpectationEnvironment(self):
        s = System([strip_all_but_identity(c) for c in self.corners], [strip_all_but_identity(x) for x in self.sides],
                   self.state, strip_all_but_identity(self.op_center), self.state_conj, rng=self.rng)
        s.solver = self.solver
        return s

    def computeOneSiteExpectation(self):
        """_2d.py:396-428."""
        expectation = 0
        stripped = self.stripExpectationEnvironment()
        for tag, value in self.op_center.items():
            if isinstance(tag, OneSiteOperator):
                s = _copy.copy(stripped)
                s.op_center[tag] = value
                expectation += s.computeExpectation()
            if isinstance(tag, TwoSiteOperator):
                if tag.position != 0:
                    continue
                for dir_, other_dir in ((0, 2), (1, 3)):
                    if tag.direction == dir_:
                        other = tag.with_(other_dir, 0)
                        s = _copy.copy(stripped)
                        s.op_center[tag] = value
                        s.op_center[other] = self.op_center[other]
                        s.contractTowards(dir_)
                        expectation += s.computeExpectation()
        return expectation

    def computeEstimatedOneSiteExpectation(self, direction=0):
        """base.py:58-64."""
        s = _copy.copy(self)
        e1 = s.computeExpectation()
        s.contractTowards(direction)
        return s.computeExpectation() - e1

    # -- the solve (_2d.py:489-502) ---------------------------------------------------
    def minimizeExpectation(self, max_mults=100, tol=1e-7):
        op = self.formOperator()
        k = 3
        use_lu = solver.is_cheaper_to_form_matrix(10 * 2 * k, op.cost_N, op.cost_N_form, op.n)
        use_dense_H = solver.is_cheaper_to_form_matrix(2 * k, op.cost_H, op.cost_H_form, op.n)
        stats = {}
        v = solver.relax_over(self.state.ravel(), lambda x: op.apply_H(x).ravel(),
                              lambda x: op.apply_N(x).ravel(), max_mults=max_mults, tol=tol,
                              H_matrix=op.matrix_H() if use_dense_H else None,
                              N_matrix=op.matrix_N() if use_lu else None,
                              use_dense_H=use_dense_H, use_lu_N=use_lu, stats=stats)
        self.last_solve = dict(stats, use_lu=use_lu, use_dense_H=use_dense_H)
        self.setStateCenter(v.reshape(self.state.shape))

    def minimizeExpectationUsingFullEigensolver(self):
        op = self.formOperator()
        w, V = eigh(op.matrix_H(), op.matrix_N())
        self.setStateCenter(V[:, 0].reshape(self.state.shape))

    # -- gauge moves (_2d.py:503-549) --------------------------------------------------
    def normalize(self):
        for c in range(4):
            for d in range(2):
                self.normalizeCornerAndDenormalizeSide(c, d)
        for s in range(4):
            self.normalizeSideAndDenormalizeCenter(s)

    def normalizeCenterAndDenormalizeSide(self, direction):
        nc, ds1 = linalg.normalize_axis(self.state, direction, True)
        self.state = linalg.absorb_matrix_at(self.state, direction, nc)
        self.state_conj = self.state.conj()
        ds2 = ds1.conj()
        self.sides[direction] = _map(lambda t: linalg.absorb_matrix_at(linalg.absorb_matrix_at(t, 6, ds1), 7, ds2), self.sides[direction])

    def normalizeCornerAndDenormalizeSide(self, corner_id, direction):
        side_id = side_from_corner(corner_id, direction)
        ca1, ca2 = direction * 3, direction * 3 + 1
        sa1, sa2 = (1 - direction) * 3, (1 - direction) * 3 + 1
        n1, d1 = linalg.normalize_axis(self.corners[corner_id][Identity()], ca1, True)
        n2, d2 = n1.conj(), d1.conj()
        self.corners[corner_id] = _map(lambda t: linalg.absorb_matrix_at(linalg.absorb_matrix_at(t, ca1, n1), ca2, n2), self.corners[corner_id])
        self.sides[side_id] = _map(lambda t: linalg.absorb_matrix_at(linalg.absorb_matrix_at(t, sa1, d1), sa2, d2), self.sides[side_id])

    def normalizeSideAndDenormalizeCenter(self, side_id):
        n1, dc = linalg.normalize_axis(self.sides[side_id][Identity()], 6, True)
        n2 = n1.conj()
        self.sides[side_id] = _map(lambda t: linalg.absorb_matrix_at(linalg.absorb_matrix_at(t, 6, n1), 7, n2), self.sides[side_id])
        self.state = linalg.absorb_matrix_at(self.state, side_id, dc)
        self.state_conj = self.state.conj()

    # -- bandwidth increase (base.py:114-160, _2d.py:460-467) ---------------------------
    def increaseBandwidth(self, direction, by=None, to=None, do_as_much_as_possible=False, enlargeners=None):
        if direction not in (0, 1):
            raise ValueError("Direction for bandwidth increase must be either 0 (for horizontal axes) "
                             "or 1 (for vertical axes), not {}.".format(direction))
        axis = direction
        S = self.state
        O_axis = O(axis)
        d = S.shape[-1]
        old = S.shape[axis]
        if (by is None) == (to is None):
            raise ValueError("Either 'by' or 'to' must be given, not both.")
        new = old + by if by is not None else to
        assert new >= old
        if new == old:
            return
        if new > d * old:
            if do_as_much_as_possible:
                new = d * old
            else:
                raise ValueError("New dimension must be less than the physical dimension times the old dimension "
                                 "({} > {}*{}).".format(new, d, old))
        n0 = linalg.normalize_axis(S, O_axis)[0]
        n1 = linalg.normalize_axis(S, axis)[0]
        if enlargeners is None:
            eA, eB = linalg.new_enlargener(old, new, self.rng)
        else:
            eA, eB = enlargeners
        S = linalg.absorb_matrix_at(S, axis, eA)
        n0 = linalg.absorb_matrix_at(n0, O_axis, eB)
        n0, S = linalg.normalize_axis_and_denormalize(n0, O_axis, axis, S)
        n1 = linalg.absorb_matrix_at(n1, axis, eA)
        S = linalg.absorb_matrix_at(S, O_axis, eB)
        n1, S = linalg.normalize_axis_and_denormalize(n1, axis, O_axis, S)
        self.setStateCenter(S)
        self.contractUnnormalizedTowards(axis, n0)
        self.contractUnnormalizedTowards(O_axis, n1)
        self.just_increased_bandwidth = True
        return eA, eB

    # -- driver (base.py:65-111) -------------------------------------------------------
    def setPolicy(self, name, policy):
        if name not in self.policies:
            raise ValueError("No such policy name " + name)
        if self.policies[name] is not None:
            raise ValueError("Policy " + name + " has already been set.")
        self.policies[name] = policy.bind(self)

    def _policy(self, name, optional=False):
        p = self.policies.get(name, KeyError)
        if p is KeyError:
            raise ValueError("No such policy name " + name)
        if p is None:
            if optional:
                return []
            raise ValueError("Policy " + name + " has not been set.")
        return [p]

    def sweepUntilConverged(self, max_steps=None):
        self.number_of_sweeps = getattr(self, "number_of_sweeps", 0) + 1
        for p in self._policy("contraction"):
            p.reset()
        for p in self._policy("sweep convergence"):
            p.reset()
        for p in self._policy("pre-optimization hook", True):
            p.apply()
        self.minimizeExpectation()
        for p in self._policy("post-optimization hook", True):
            p.apply()
        for p in self._policy("sweep convergence"):
            p.update()
        steps = 0
        while not self._policy("sweep convergence")[0].converged():
            if max_steps is not None and steps >= max_steps:
                break
            steps += 1
            for name in ("contraction",):
                for p in self._policy(name):
                    p.apply()
            for name in ("state compression", "operator compression"):
                for p in self._policy(name, True):
                    p.apply()
            self.number_of_iterations = getattr(self, "number_of_iterations", 0) + 1
            try:
                for p in self._policy("pre-optimization hook", True):
                    p.apply()
                self.minimizeExpectation()
                for p in self._policy("post-optimization hook", True):
                    p.apply()
                for p in self._policy("sweep convergence"):
                    p.update()
            except solver.RelaxFailed:
                pass

    def runUntilConverged(self, max_runs=None, max_steps=None):
        self.number_of_sweeps = 0
        self.number_of_iterations = 0
        self.sweepUntilConverged(max_steps)
        for p in self._policy("run convergence"):
            p.update()
        runs = 0
        while not self._policy("run convergence")[0].converged():
            if max_runs is not None and runs >= max_runs:
                break
            runs += 1
            for p in self._policy("bandwidth increase"):
                p.apply()
            self.sweepUntilConverged(max_steps)
            for p in self._policy("run convergence"):
                p.update()
