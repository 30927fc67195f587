"""One ground-state solve of a device-resident environment (the unit of the benchmark).

solve(...) = System.minimizeExpectation's work (carcassonne/system/_2d.py:489-497) with the
residual-controlled FP64 Lanczos in place of relaxOver: stage 1 and 2 on the GPU, the
stage-3 operators with packed slabs, the normalization factor, the eigensolve, and a final
FP64 Rayleigh quotient.
"""
import time

import numpy as np
import torch

from .data import HipData, PermutedView
from .precision import MixedSolver
from .solver import GroundStateSolver
from .tensors._2d import sparse as S2


def synthetic_tag(k):
    """The tag of a synthetic environment's plain-tuple channel key: ("I",) Identity, ("C",) Complete,
    ("O",) OneSiteOperator, ("T", k) the side's open TwoSiteOperator(k, CENTER) channel, ("TT", k, dir) the
    center operator's two-site halves."""
    from .sparse import Identity, Complete, OneSiteOperator, TwoSiteOperator, CENTER
    if k[0] == "I":
        return Identity()
    if k[0] == "C":
        return Complete()
    if k[0] == "O":
        return OneSiteOperator(None)
    if k[0] == "T":
        return TwoSiteOperator(k[1], CENTER)
    return TwoSiteOperator(k[1], k[2], 0)


def device_environment(corners, sides, op_center):
    """Upload a synthetic environment (carcassonne_amd.synthetic: plain-tuple channel keys, see
    synthetic_tag) as tag -> HipData dicts."""
    tag = synthetic_tag
    up = HipData.fromArray
    return ([{tag(k): up(v) for k, v in c.items()} for c in corners],
            [{tag(k): up(v) for k, v in s.items()} for s in sides],
            {tag(k): up(v) for k, v in op_center.items()})


class Stage12Graph:
    """Stage 1 and 2 of the expectation multipliers (sparse.py:59-71 without stage 3) recorded once as a
    HIP graph for one environment structure and replayed per solve.  The Python bookkeeping of the
    sparse contractions (tag rules, contraction plans, tensor objects) runs only at capture; a replay
    recomputes every stage-1/2 tensor in place from the graph's static inputs, which run() first fills
    from the caller's environment.  Stage 3 stays eager, so every solve builds fresh operators; the
    slab-channel caches attached to the stage-2 tensors are dropped and lazily permuted results are
    returned to their lazy state after each replay.  The stage-2 tensors of one replay are valid until
    the next run() of the same graph."""

    def __init__(self, corners, sides):
        self._static = ([{k: HipData(v._t.clone()) for k, v in c.items()} for c in corners],
                        [{k: HipData(v._t.clone()) for k, v in s.items()} for s in sides])
        sc, ss = self._static
        main = torch.cuda.current_stream()
        warm = torch.cuda.Stream()
        warm.wait_stream(main)
        with torch.cuda.stream(warm):  # first launches of every kernel happen outside the capture
            self._stages(sc, ss)
        main.wait_stream(warm)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.stage2 = self._stages(sc, ss)
        self._lazy = [(t, t._base, t._perm) for b in self.stage2 for t in b.values()
                      if isinstance(t, PermutedView) and t._lazy()]

    @staticmethod
    def _stages(corners, sides):
        st1 = [S2.formExpectationStage1(corners[i], sides[i]) for i in range(4)]
        return S2.formExpectationStage2(st1[0], st1[1]), S2.formExpectationStage2(st1[2], st1[3])

    @staticmethod
    def key(corners, sides):
        return tuple(tuple((k, tuple(v.shape)) for k, v in d.items()) for d in list(corners) + list(sides))

    def run(self, corners, sides):
        sc, ss = self._static
        dsts, srcs = [], []
        for dst, src in zip(sc + ss, list(corners) + list(sides)):
            for k, v in src.items():
                dsts.append(dst[k]._t)
                srcs.append(v._t)
        torch._foreach_copy_(dsts, srcs)   # one multi-tensor launch for the whole environment
        self.graph.replay()
        for b in self.stage2:
            for t in b.values():
                t.__dict__.pop("_stage3_slabs", None)
        for t, base, perm in self._lazy:
            t._base, t._perm, t._mat = base, perm, None
        return self.stage2


_STAGE12 = {}


def multipliers(corners, sides, op_center, graph=True):
    """(H, N) of formExpectationAndNormalizationMultipliers, with stage 1/2 from a replayed Stage12Graph
    (one per environment structure; at most four kept) unless graph is False."""
    if not graph:
        return S2.formExpectationAndNormalizationMultipliers(corners, sides, op_center)
    key = Stage12Graph.key(corners, sides)
    g = _STAGE12.get(key)
    if g is None:
        if len(_STAGE12) >= 4:
            _STAGE12.clear()
        g = _STAGE12[key] = Stage12Graph(corners, sides)
    b0, b1 = g.run(corners, sides)
    return S2.formExpectationStage3(b0, b1, op_center)


def flops_per_matvec(operator):
    """Algorithmic real FLOPs of one application of a stage-3 operator: per term
    4*chi^2*D^6*d (two GEMMs over the environment) + 2*D^4*d^2 when the term has a site
    operator -- the reference's CostTracker count converted to real FLOPs."""
    f = 0
    for A, B, O in zip(operator._A, operator._B, operator._O):
        na, P = operator.n_alpha, operator.P
        f += 2 * (na * P * operator.Q * operator.d * P) + 2 * (operator.Q * P * operator.d * operator.Q * na)
        if O is not None:
            f += 2 * operator.d * P * operator.Q * operator.d
    return f


def bytes_per_matvec(operator, element_bytes=8):
    """Algorithmic HBM bytes of one fused application: every distinct slab channel once
    (chi^2 * P * P elements each) plus the input and output vectors."""
    na, P = operator.n_alpha, operator.P
    a = len({id(x) for x in operator._A})
    b = len({id(x) for x in operator._B})
    return (a + b) * na * P * operator.Q * element_bytes + 2 * operator.n * 8


def solve(corners, sides, op_center, x0, tol_residual=1e-10, tol_energy=1e-13, max_iter=400, check_every=4,
          graph=True):
    """The plain FP64 form: returns x (N-normalised), the Ritz value, the energy (FP64 Rayleigh quotient),
    the residual estimate and the Lanczos iteration count."""
    H, N = multipliers(corners, sides, op_center, graph=graph)
    solver = GroundStateSolver(H.operator, N.operator)
    x, theta, resid, iters = solver.solve(x0, tol_residual=tol_residual, tol_energy=tol_energy,
                                          max_iter=max_iter, check_every=check_every)
    energy = solver.rayleigh_quotient(x)
    return dict(x=x, theta=theta, energy=energy, residual=resid, iterations=iters, operator=H.operator)


def solve_mixed(corners, sides, op_center, x0, dtype="f16", graph=True, **kw):
    """The precision form of the same solve: stage 1/2 and the normalization factor in FP64, FP16/BF16
    Lanczos on the whitened operator, FP64 Jacobi-Davidson refinement.  The energy is the FP64
    Rayleigh-Ritz value of the refined vector (its Rayleigh quotient, from FP64 H and N products)."""
    H, N = multipliers(corners, sides, op_center, graph=graph)
    solver = MixedSolver(H.operator, N.operator, dtype=dtype)
    x, rho, info = solver.solve(x0, **kw)
    return dict(x=x, energy=rho, iterations=info["lowprec_iterations"], info=info, operator=H.operator,
                lp_operator=solver.Hlp)
