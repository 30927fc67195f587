"""The 1D <-> 2D consistency harness (carcassonne/system/_1d2d.py), test infrastructure: a 2D system on the
GPU (carcassonne_amd.system.System, data on the device) and the NumPy 1D system (oracle/system1d.py) of the
same chain are stepped together, and after every step the center states (up to a phase and to a
threshold), the environments, the expectation, the expectation matrix and the 2D normalization are
compared, as the reference's harness does (_1d2d.py:64-115).  The class is itself a sweep system, so the
policies of carcassonne_amd.policies drive it (BaseSystem).  As in the reference, the bandwidth increase
hands the 2D side's random enlargeners to the 1D side, and both sides take their polar factors from one
SVD routine (see _device_normalize_axis)."""
from copy import copy

import numpy as np

from carcassonne_amd.data import HipData
from carcassonne_amd.system import System as System2D
from carcassonne_amd.system.base import BaseSystem
from oracle.system1d import System1D, make_mpo
from tests.oracle_bridge import oracle_to_product_tag


class ConsistencyError(Exception):
    pass


def _device_normalize_axis(a, axis):
    """normalizeAxis of a host array by the 2D system's own routine (HipData.normalizeAxis, the device SVD).
    The reference's harness runs both systems on one SVD routine; the polar factor of a rank-deficient
    matricization (the near-product states of these chains) is not unique -- the singular vectors of zero
    singular values are each SVD's own choice -- so the 1D side takes its polar factors from the same
    routine as the 2D side, and the comparison checks everything else: the MPO against the sparse tag
    algebra, the absorptions, the bandwidth increase and the minimization."""
    return tuple(x.toArray() for x in HipData.fromArray(np.ascontiguousarray(a)).normalizeAxis(axis))


class System1D2D(BaseSystem):
    @classmethod
    def new(cls, rotation, Os=(), OOs=(), adaptive_state_threshold=False, minimizer="relaxover"):
        """_1d2d.py:15-40 (real site operators: the GPU path is real)."""
        I = np.identity(2)
        operator, right_boundary, right_tags, left_boundary, left_tags = make_mpo(I, Os, OOs)
        one_d = System1D(right_boundary, left_boundary, operator, np.ones((1, 1, 2)) / np.sqrt(2))
        one_d.normalize_axis = _device_normalize_axis
        Os_d = tuple(HipData.fromArray(np.asarray(O, dtype=float)) for O in Os)
        OOs_d = tuple(tuple(HipData.fromArray(np.asarray(x, dtype=float)) for x in OO) for OO in OOs)
        if rotation == 0:
            two_d = System2D.newTrivialWithSparseOperator(Os=Os_d, OO_LRs=OOs_d)
        elif rotation == 1:
            two_d = System2D.newTrivialWithSparseOperator(Os=Os_d, OO_UDs=tuple((x[1], x[0]) for x in OOs_d))
        else:
            raise ValueError("rotation must be 0 or 1, not {}".format(rotation))
        system = cls(rotation, one_d, two_d, [oracle_to_product_tag(t) for t in right_tags],
                     [oracle_to_product_tag(t) for t in left_tags], adaptive_state_threshold)
        system.minimizer = minimizer
        return system

    @classmethod
    def newSimple(cls, rotation, O=None, OO=None, adaptive_state_threshold=False, minimizer="relaxover"):
        return cls.new(rotation, [] if O is None else [O], [] if OO is None else [OO], adaptive_state_threshold,
                       minimizer)

    def __init__(self, rotation, one_d, two_d, right_tags, left_tags, adaptive_state_threshold):
        BaseSystem.__init__(self)
        self.rotation = rotation
        self._1d, self._2d = one_d, two_d
        self.right_tags, self.left_tags = right_tags, left_tags
        self.adaptive_state_threshold = adaptive_state_threshold
        self.state_threshold = 1e-5
        self.checks = 0
        self.degenerate_minimizations = []
        # "relaxover" (the reference harness's solver on both sides) or "full": the exact lowest eigenvector on
        # both sides (1D: eigh of the formed matrix, _1d.py:170-173; 2D: minimizeExpectationUsingFullEigensolver).
        # relaxOver's Krylov space cannot leave the symmetry sector of its start vector except through rounding,
        # so two implementations can end in different sectors where the start lies in an excited one.
        self.minimizer = "relaxover"

    def __copy__(self):
        other = System1D2D(self.rotation, copy(self._1d), copy(self._2d), self.right_tags, self.left_tags,
                           self.adaptive_state_threshold)
        other.minimizer = self.minimizer
        return other

    # -- checks (_1d2d.py:64-115) -----------------------------------------------------------------------
    def check(self, prefix):
        self.checkStates(prefix)
        self.checkEnvironments(prefix)
        self.checkNormalization(prefix)
        self.checkExpectation(prefix)
        self.checkExpectationMatrix(prefix)
        self.checks += 1

    def checkEnvironments(self, prefix):
        for name, one, two in (("right", self._1d.right_environment, self.convert2DRightEnvironment()),
                               ("left", self._1d.left_environment, self.convert2DLeftEnvironment())):
            err = np.linalg.norm(one - two)
            if err > 1e-7:
                raise ConsistencyError("{}: for the {} environment, norm(_1d-_2d)={} > 1e-7".format(prefix, name, err))

    def checkExpectation(self, prefix):
        e1, e2 = self._1d.computeExpectation(), self._2d.computeExpectation()
        if abs(e1 - e2) > 1e-7:
            raise ConsistencyError("{}: for the expectation, abs(1D@{} - 2D@{}) > 1e-7".format(prefix, e1, e2))

    def checkExpectationMatrix(self, prefix):
        one = self._1d.formExpectationMatrix()
        two = self._2d.formExpectationMatrix().toArray().reshape(one.shape)
        err = np.linalg.norm(one - two)
        if err > 1e-7:
            raise ConsistencyError("{}: for the expectation matrix, norm(1D - 2D)={} > 1e-7".format(prefix, err))

    def checkNormalization(self, prefix):
        n = self._2d.computeNormalization()
        if abs(n - 1) > 1e-7:
            raise ConsistencyError("{}: the 2D normalization is {} != 1".format(prefix, n))

    @staticmethod
    def _phase_fixed(a):
        a = np.array(a, dtype=float).copy()
        for x in a.ravel():
            if abs(x) > 1e-12:
                return a / x
        return a

    def checkStates(self, prefix):
        one = self._phase_fixed(self._1d.state_center_data)
        two = self._phase_fixed(self._2d.state_center_data.toArray().reshape(one.shape))
        rel = np.linalg.norm(one - two) / np.linalg.norm(np.abs(one) + np.abs(two)) * 2
        if rel > self.state_threshold:
            raise ConsistencyError("{}: for the center state, norm(_1d-_2d)={} > {}".format(prefix, rel, self.state_threshold))
        return self._2d.state_center_data

    # -- environment conversions and copies (_1d2d.py:128-167) ------------------------------------------
    def _convert2D(self, side_id, tags):
        b = self._2d.state_center_data.shape[self.rotation]
        out = np.zeros((len(tags), b, b))
        for tag, value in self._2d.sides[side_id].items():
            out[tags.index(tag)] = value.toArray().reshape(b, b)
        return out

    def convert2DLeftEnvironment(self):
        return self._convert2D(self.rotation + 2, self.left_tags)

    def convert2DRightEnvironment(self):
        return self._convert2D(self.rotation, self.right_tags)

    def copy2Dto1D(self):
        s = self._2d.state_center_data.toArray()
        self._1d.setStateCenter(s.reshape(s.shape[0] * s.shape[1], s.shape[2] * s.shape[3], s.shape[4]))
        self._1d.left_environment = self.convert2DLeftEnvironment()
        self._1d.right_environment = self.convert2DRightEnvironment()

    # -- the steps (_1d2d.py:116-127, 168-203) ------------------------------------------------------------
    @property
    def state_center_data(self):
        return self.checkStates("when fetching state")

    def setStateCenter(self, state):
        self._2d.setStateCenter(state)

    def computeExpectation(self):
        self.check("while computing expectation")
        return self._2d.computeExpectation()

    def computeOneSiteExpectation(self):
        one, two = self._1d.computeOneSiteExpectation(), self._2d.computeOneSiteExpectation()
        if abs(abs(one) - abs(two)) > 1e-7:
            raise ConsistencyError("for the one-site expectation, abs({}-{}) > 1e-7".format(abs(one), abs(two)))
        return two

    def contractTowards(self, direction):
        self.check("before contraction")
        self._1d.contractTowards(direction)
        self._2d.contractTowards(2 * direction + self.rotation)
        self.check("after contraction")

    def increaseBandwidth(self, direction, by=None, to=None, do_as_much_as_possible=False):
        self.check("before increasing bandwidth")
        assert direction == 0
        enlargeners = self._2d.increaseBandwidth(self.rotation, by, to, do_as_much_as_possible)
        if enlargeners is not None:
            enlargeners = tuple(e.toArray() if hasattr(e, "toArray") else np.asarray(e) for e in enlargeners)
        self._1d.increaseBandwidth(0, by, to, do_as_much_as_possible, enlargeners=enlargeners)
        self.check("after increasing bandwidth")

    def minimizeExpectation(self):
        self.check("before minimizing")
        self.copy2Dto1D()
        original_threshold = self.state_threshold
        if self.adaptive_state_threshold:
            # as _1d2d.py:183-194: each vector divided in turn by every entry above 1e-12 (in place, so each
            # division uses the entry's current value)
            def scaled(v):
                v = np.array(v, dtype=float).ravel().copy()
                for i in range(v.size):
                    if abs(v[i]) > 1e-12:
                        v /= v[i]
                return v
            original = scaled(self._1d.state_center_data)
            multiplied = scaled(self._1d.multiply(self._1d.state_center_data))
            diff = np.linalg.norm(original - multiplied) / (np.linalg.norm(original) + np.linalg.norm(multiplied)) * 2
            self.state_threshold /= diff
        if self.minimizer == "full":
            self._1d.minimizeExpectationUsingFullEigensolver()
            self._2d.minimizeExpectationUsingFullEigensolver()
        else:
            self._1d.minimizeExpectation()
            self._2d.minimizeExpectation()
        try:
            self.check("after minimizing")
        except ConsistencyError as err:
            # Where the lowest eigenvalue is (nearly) degenerate -- the Heisenberg chain's spin doublets at odd
            # site counts -- the two solves may stop at different vectors of one eigenspace: accepted when
            # both have the same energy to 1e-9 and everything but the state agrees; the 1D side then takes
            # the 2D state (copy2Dto1D below), as after every minimization.
            e1, e2 = self._1d.computeExpectation(), self._2d.computeExpectation()
            if "center state" not in str(err) or abs(e1 - e2) > 1e-9 * max(abs(e2), 1.0):
                raise ConsistencyError("{} (energies 1D {!r}, 2D {!r}; 1D state dtype {}, solves 2D {})".format(
                    err, e1, e2, self._1d.state_center_data.dtype, self._2d.last_solve)) from err
            self.degenerate_minimizations.append((e1, e2, str(err)))
            self.copy2Dto1D()
            self.check("after minimizing (degenerate)")
        self.state_threshold = original_threshold
        self.copy2Dto1D()
