"""Shape-only Data backend that counts complex multiply-adds of a contraction sequence.
Restates carcassonne/data/cost_tracker.py:8-40: each pairwise contraction costs
prod(result shape) * prod(contracted extents); joins are free."""
from math import prod

from . import Data


class CostTracker(Data):
    def __init__(self, shape, cost=0):
        self.shape = tuple(int(s) for s in shape)
        self.ndim = len(self.shape)
        self.cost = cost

    def __repr__(self):
        return "CostTracker({},{})".format(self.shape, self.cost)

    def contractWith(self, other, self_axes, other_axes):
        self_axes, other_axes = list(self_axes), list(other_axes)
        shape = tuple(x for i, x in enumerate(self.shape) if i not in self_axes) + \
            tuple(x for i, x in enumerate(other.shape) if i not in other_axes)
        cost = self.cost + other.cost + prod(shape) * prod(x for i, x in enumerate(self.shape) if i in self_axes)
        return CostTracker(shape, cost)

    def extractScalar(self):
        assert len(self.shape) == 0
        return self

    def join(self, *groups):
        groups = [[g] if isinstance(g, int) else list(g) for g in groups]
        return CostTracker(tuple(prod(self.shape[i] for i in g) for g in groups), self.cost)


def computeCostOfContracting(contractor, *tensors):
    return contractor(*(CostTracker(t.shape) if not isinstance(t, tuple) else CostTracker(t) for t in tensors)).cost


__all__ = ["CostTracker", "computeCostOfContracting"]
