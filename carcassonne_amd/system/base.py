"""Policy slots and the sweep / run driver.  Re-implements carcassonne/system/base.py:
_applyPolicy/_getPolicy/setPolicy (45-57,77-84), computeEstimatedOneSiteExpectation (58-64),
runUntilConverged (65-76), sweepUntilConverged (85-111, with optional step caps; the
reference loops until the policy fires) and _increaseBandwidth (114-160)."""
import logging
from copy import copy

from ..errors import RelaxFailed
from ..utils import O, computeNewDimension

log = logging.getLogger(__name__)


class BaseSystem:
    POLICY_NAMES = ["bandwidth increase", "contraction", "operator compression", "run convergence",
                    "post-contraction hook", "pre-optimization hook", "post-optimization hook",
                    "state compression", "sweep convergence"]

    def __init__(self):
        self._policies = {name: None for name in self.POLICY_NAMES}

    def _getPolicy(self, policy_name, optional=False):
        if policy_name not in self._policies:
            raise ValueError("No such policy name " + policy_name)
        p = self._policies[policy_name]
        if p is None:
            if optional:
                return []
            raise ValueError("Policy " + policy_name + " has not been set.")
        return [p]

    def _applyPolicy(self, policy_name, optional=False):
        for p in self._getPolicy(policy_name, optional):
            p.apply()

    def _hasConverged(self, policy_name):
        for p in self._getPolicy(policy_name, False):
            return p.converged()

    def _resetPolicy(self, policy_name, optional=False):
        for p in self._getPolicy(policy_name, optional):
            p.reset()

    def _updatePolicy(self, policy_name, optional=False):
        for p in self._getPolicy(policy_name, optional):
            p.update()

    def setPolicy(self, policy_name, policy):
        if policy_name not in self._policies:
            raise ValueError("No such policy name " + policy_name)
        if self._policies[policy_name] is not None:
            raise ValueError("Policy " + policy_name + " has already been set.")
        self._policies[policy_name] = policy.createBindingToSystem(self)

    def computeEstimatedOneSiteExpectation(self, direction=0):
        system = copy(self)
        exp1 = system.computeExpectation()
        system.contractTowards(direction)
        return system.computeExpectation() - exp1

    def runUntilConverged(self, maximum_number_of_runs=None, maximum_number_of_steps=None):
        log.info("Beginning run.")
        self.number_of_sweeps = 0
        self.number_of_iterations = 0
        self.sweepUntilConverged(maximum_number_of_steps)
        self._updatePolicy("run convergence")
        runs = 0
        while not self._hasConverged("run convergence"):
            if maximum_number_of_runs is not None and runs >= maximum_number_of_runs:
                break
            runs += 1
            self._applyPolicy("bandwidth increase")
            self.sweepUntilConverged(maximum_number_of_steps)
            self._updatePolicy("run convergence")
        log.info("Finished run with {} total sweeps and {} total iterations.".format(
            self.number_of_sweeps, self.number_of_iterations))

    def sweepUntilConverged(self, maximum_number_of_steps=None):
        self.number_of_sweeps = getattr(self, "number_of_sweeps", 0) + 1
        self._resetPolicy("contraction")
        self._resetPolicy("sweep convergence")
        self.iteration_number_for_sweep = 1
        self._applyPolicy("pre-optimization hook", optional=True)
        self.minimizeExpectation()
        self._applyPolicy("post-optimization hook", optional=True)
        self._updatePolicy("sweep convergence")
        steps = 0
        while not self._hasConverged("sweep convergence"):
            if maximum_number_of_steps is not None and steps >= maximum_number_of_steps:
                break
            steps += 1
            self._applyPolicy("contraction")
            self._applyPolicy("state compression", optional=True)
            self._applyPolicy("operator compression", optional=True)
            self.iteration_number_for_sweep += 1
            self.number_of_iterations = getattr(self, "number_of_iterations", 0) + 1
            try:
                self._applyPolicy("pre-optimization hook", optional=True)
                self.minimizeExpectation()
                self._applyPolicy("post-optimization hook", optional=True)
                self._updatePolicy("sweep convergence")
            except RelaxFailed:
                pass

    def _increaseBandwidth(self, axis, by=None, to=None, do_as_much_as_possible=False, enlargeners=None):
        state = self.state_center_data
        O_axis = O(axis) if state.ndim == 5 else 1 - axis
        physical_dimension = state.shape[-1]
        old_dimension = state.shape[axis]
        new_dimension = computeNewDimension(old_dimension, by=by, to=to)
        if new_dimension == old_dimension:
            return
        if new_dimension > physical_dimension * old_dimension:
            if do_as_much_as_possible:
                new_dimension = physical_dimension * old_dimension
            else:
                raise ValueError("New dimension must be less than the physical dimension times the old dimension "
                                 "({} > {}*{}).".format(new_dimension, physical_dimension, old_dimension))
        neighbor_0 = state.normalizeAxis(O_axis)[0]
        neighbor_1 = state.normalizeAxis(axis)[0]
        if enlargeners is None:
            enlargener_A, enlargener_B = state.newEnlargener(old_dimension, new_dimension)
        else:
            enlargener_A, enlargener_B = enlargeners
        state = state.absorbMatrixAt(axis, enlargener_A)
        neighbor_0 = neighbor_0.absorbMatrixAt(O_axis, enlargener_B)
        neighbor_0, state = neighbor_0.normalizeAxisAndDenormalize(O_axis, axis, state)
        neighbor_1 = neighbor_1.absorbMatrixAt(axis, enlargener_A)
        state = state.absorbMatrixAt(O_axis, enlargener_B)
        neighbor_1, state = neighbor_1.normalizeAxisAndDenormalize(axis, O_axis, state)
        self.setStateCenter(state)
        self.contractUnnormalizedTowards(axis, neighbor_0)
        self.contractUnnormalizedTowards(O_axis, neighbor_1)
        self.just_increased_bandwidth = True
        return enlargener_A, enlargener_B
