"""Strategy objects for the driver's named slots.  Re-implements carcassonne/policies.py:12-242
(policy/proxy binding, bandwidth increase, state compression, contraction pattern,
convergence tests, hooks) with the same names and behaviour."""
import logging

import numpy as np

from .utils import O

log = logging.getLogger(__name__)


class Policy:
    def createBindingToSystem(self, system):
        return self.Proxy(self, system)


class Proxy:
    def __init__(self, forward, system):
        self.forward = forward
        self.system = system

    def __getattr__(self, name):
        return getattr(self.forward, name)


class ApplyProxy(Proxy):
    def apply(self):
        return type(self.forward).apply(self)


class ConvergedProxy(Proxy):
    def converged(self):
        return type(self.forward).converged(self)


class ResetProxy(Proxy):
    def reset(self):
        return type(self.forward).reset(self)


class UpdateProxy(Proxy):
    def update(self):
        return type(self.forward).update(self)


class _StatefulProxy(ConvergedProxy, ResetProxy, UpdateProxy):
    """Convergence proxies keep their own last/current values per bound system."""

    def __setattr__(self, name, value):
        object.__setattr__(self, name, value)


class BandwidthIncreasePolicy(Policy):
    Proxy = ApplyProxy


class AlternatingDirectionsIncrementBandwidthIncreasePolicy(BandwidthIncreasePolicy):
    def __init__(self, directions, increment=1):
        self.directions, self.direction, self.increment = directions, 0, increment

    def apply(self):
        self.system.increaseBandwidth(self.direction, by=self.increment, do_as_much_as_possible=True)
        self.direction = O(self.direction)


class AllDirectionsIncrementBandwidthIncreasePolicy(BandwidthIncreasePolicy):
    def __init__(self, increment=1):
        self.increment = increment

    def apply(self):
        for direction in (0, 1):
            self.system.increaseBandwidth(direction, by=self.increment, do_as_much_as_possible=True)


class OneDirectionIncrementBandwidthIncreasePolicy(BandwidthIncreasePolicy):
    def __init__(self, direction, increment=1):
        self.direction, self.increment = direction, increment

    def apply(self):
        self.system.increaseBandwidth(self.direction, by=self.increment, do_as_much_as_possible=True)


class CompressionPolicy(Policy):
    Proxy = ApplyProxy


class ConstantStateCompressionPolicy(CompressionPolicy):
    def __init__(self, new_dimension):
        self.new_dimension = new_dimension

    def apply(self):
        for corner_id in range(4):
            for direction in range(2):
                self.system.compressCornerStateTowards(corner_id, direction, self.new_dimension)


class ContractionPolicy(Policy):
    class Proxy(ApplyProxy, ResetProxy):
        pass


class RepeatPatternContractionPolicy(ContractionPolicy):
    def __init__(self, directions):
        self.directions = list(directions)
        self.iteration = iter([])

    def apply(self):
        try:
            direction = next(self.iteration)
        except StopIteration:
            self.iteration = iter(self.directions)
            try:
                direction = next(self.iteration)
            except StopIteration:
                raise ValueError("An empty sequence of contraction directions was provided! ({})".format(self.directions))
        self.system.contractTowards(direction)

    def reset(self):
        self.iteration = iter([])


class ConvergencePolicy(Policy):
    Proxy = _StatefulProxy


class PeriodicyThresholdConvergencePolicy(ConvergencePolicy):
    def __init__(self, threshold, *directions):
        if not directions:
            raise ValueError("at least one direction must be specified")
        self.threshold, self.directions = threshold, directions

    def converged(self):
        state = self.system.state_center_data
        difference = 0.0
        for direction in self.directions:
            normalized = state.normalizeAxis(direction)[0]
            denormalizer = state.normalizeAxis(O(direction))[-1]
            difference += (state - normalized.absorbMatrixAt(direction, denormalizer)).norm()
        return difference < self.threshold

    def reset(self):
        pass

    def update(self):
        pass


class RelativeEstimatedOneSiteExpectationDifferenceThresholdConvergencePolicy(ConvergencePolicy):
    def __init__(self, threshold, direction=0):
        self.threshold, self.direction = threshold, direction
        self.last = self.current = None

    def converged(self):
        last, current = self.last, self.current
        if last is not None and current is not None:
            magnitude = abs(current + last)
            return magnitude < 1e-15 or abs(current - last) / magnitude * 2 < self.threshold

    def reset(self):
        self.last = self.current = None

    def update(self):
        self.last = self.current
        self.current = self.system.computeEstimatedOneSiteExpectation(self.direction)


class RelativeExpectationDifferenceDifferenceThresholdConvergencePolicy(ConvergencePolicy):
    def __init__(self, threshold):
        self.threshold = threshold
        self.last_value = self.last_difference = self.current_value = self.current_difference = None

    def converged(self):
        last, current = self.last_difference, self.current_difference
        if last is not None and current is not None:
            absolute = abs(current - last)
            return absolute < 1e-15 or absolute / abs(current + last) * 2 < self.threshold

    def reset(self):
        self.last_value = self.last_difference = self.current_value = self.current_difference = None

    def update(self):
        self.last_value, self.last_difference = self.current_value, self.current_difference
        self.current_value = self.system.computeExpectation()
        if self.last_value is not None:
            self.current_difference = self.current_value - self.last_value


class RelativeOneSiteExpectationDifferenceThresholdConvergencePolicy(ConvergencePolicy):
    def __init__(self, threshold):
        self.threshold = threshold
        self.last = self.current = None

    def converged(self):
        last, current = self.last, self.current
        if last is not None and current is not None:
            absolute = abs(current - last)
            return absolute < 1e-15 or absolute / abs(current + last) * 2 < self.threshold

    def reset(self):
        self.last = self.current = None

    def update(self):
        expectation = self.system.computeOneSiteExpectation()
        self.last, self.current = self.current, expectation


class RelativeStateDifferenceThresholdConvergencePolicy(ConvergencePolicy):
    def __init__(self, threshold):
        self.threshold = threshold
        self.last = self.current = None

    def converged(self):
        last, current = self.last, self.current
        if last is not None and current is not None and last.shape == current.shape:
            magnitude = (current + last).norm()
            return magnitude < 1e-15 or (current - last).norm() / magnitude * 2 < self.threshold
        return False

    def reset(self):
        self.last = self.current = None

    def update(self):
        self.last = self.current
        self.current = self.system.state_center_data.copy()


class HookPolicy(Policy):
    Proxy = ApplyProxy

    def __init__(self, callback):
        self.callback = callback

    def apply(self):
        return self.callback(self.system)


__all__ = [n for n in dir() if n.endswith("Policy") or n.endswith("Proxy") or n == "Policy"]
