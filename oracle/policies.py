"""Sweep policies (oracle).  Restates carcassonne/policies.py:46-242.

A policy object is bound to a system with ``bind`` (the reference's Proxy,
policies.py:12-43); the bound copy holds the per-system state.
"""
import copy as _copy

import numpy as np

from . import linalg
from .tags import O


class Policy:
    def bind(self, system):
        b = _copy.copy(self)
        b.system = system
        return b


class OneDirectionIncrementBandwidthIncreasePolicy(Policy):
    def __init__(self, direction, increment=1):
        self.direction, self.increment = direction, increment

    def apply(self):
        self.system.increaseBandwidth(self.direction, by=self.increment, do_as_much_as_possible=True)


class AllDirectionsIncrementBandwidthIncreasePolicy(Policy):
    def __init__(self, increment=1):
        self.increment = increment

    def apply(self):
        for direction in (0, 1):
            self.system.increaseBandwidth(direction, by=self.increment, do_as_much_as_possible=True)


class ConstantStateCompressionPolicy(Policy):
    def __init__(self, new_dimension):
        self.new_dimension = new_dimension

    def apply(self):
        for corner_id in range(4):
            for direction in range(2):
                self.system.compressCornerStateTowards(corner_id, direction, self.new_dimension)


class RepeatPatternContractionPolicy(Policy):
    def __init__(self, directions):
        self.directions = list(directions)
        self.reset()

    def apply(self):
        try:
            direction = next(self.iteration)
        except StopIteration:
            self.iteration = iter(self.directions)
            try:
                direction = next(self.iteration)
            except StopIteration:
                raise ValueError("An empty sequence of contraction directions was provided!")
        self.system.contractTowards(direction)

    def reset(self):
        self.iteration = iter([])


class _Convergence(Policy):
    def reset(self):
        self.last = self.current = None


class RelativeOneSiteExpectationDifferenceThresholdConvergencePolicy(_Convergence):
    def __init__(self, threshold):
        self.threshold = threshold
        self.last = self.current = None

    def converged(self):
        if self.last is not None and self.current is not None:
            a = abs(self.current - self.last)
            return a < 1e-15 or a / abs(self.current + self.last) * 2 < self.threshold
        return None

    def update(self):
        e = self.system.computeOneSiteExpectation()
        self.last, self.current = self.current, e


class RelativeEstimatedOneSiteExpectationDifferenceThresholdConvergencePolicy(_Convergence):
    def __init__(self, threshold, direction=0):
        self.threshold, self.direction = threshold, direction
        self.last = self.current = None

    def converged(self):
        if self.last is not None and self.current is not None:
            m = abs(self.current + self.last)
            return m < 1e-15 or abs(self.current - self.last) / m * 2 < self.threshold
        return None

    def update(self):
        self.last = self.current
        self.current = self.system.computeEstimatedOneSiteExpectation(self.direction)


class RelativeStateDifferenceThresholdConvergencePolicy(_Convergence):
    def __init__(self, threshold):
        self.threshold = threshold
        self.last = self.current = None

    def converged(self):
        if self.last is not None and self.current is not None:
            if self.last.shape != self.current.shape:
                return False
            m = np.linalg.norm(self.current + self.last)
            return m < 1e-15 or np.linalg.norm(self.current - self.last) / m * 2 < self.threshold
        return False

    def update(self):
        self.last = self.current
        self.current = self.system.state.copy()


class PeriodicyThresholdConvergencePolicy(_Convergence):
    def __init__(self, threshold, *directions):
        if not directions:
            raise ValueError("at least one direction must be specified")
        self.threshold, self.directions = threshold, directions

    def converged(self):
        S = self.system.state
        diff = 0
        for d in self.directions:
            iso = linalg.normalize_axis(S, d)[0]
            den = linalg.normalize_axis(S, O(d))[-1]
            diff += np.linalg.norm(S - linalg.absorb_matrix_at(iso, d, den))
        return diff < self.threshold

    def reset(self):
        pass

    def update(self):
        pass


class HookPolicy(Policy):
    def __init__(self, callback):
        self.callback = callback

    def apply(self):
        return self.callback(self.system)
