"""One ground-state solve of a device-resident environment (the unit of the benchmark).

solve(...) = System.minimizeExpectation's work (carcassonne/system/_2d.py:489-497) with the
residual-controlled FP64 Lanczos in place of relaxOver: stage 1 and 2 on the GPU, the
stage-3 operators with packed slabs, the normalization factor, the eigensolve, and a final
FP64 Rayleigh quotient.
"""
import time

import numpy as np
import torch

from .data import HipData
from .precision import MixedSolver
from .solver import GroundStateSolver
from .tensors._2d import sparse as S2


def device_environment(corners, sides, op_center):
    """Upload a synthetic environment (carcassonne_amd.synthetic: plain-tuple channel keys) as tag ->
    HipData dicts: ("I",) Identity, ("C",) Complete, ("O",) OneSiteOperator, ("T", k) the side's open
    TwoSiteOperator(k, CENTER) channel, ("TT", k, dir) the center operator's two-site halves."""
    from .sparse import Identity, Complete, OneSiteOperator, TwoSiteOperator, CENTER

    def tag(k):
        if k[0] == "I":
            return Identity()
        if k[0] == "C":
            return Complete()
        if k[0] == "O":
            return OneSiteOperator(None)
        if k[0] == "T":
            return TwoSiteOperator(k[1], CENTER)
        return TwoSiteOperator(k[1], k[2], 0)

    up = HipData.fromArray
    return ([{tag(k): up(v) for k, v in c.items()} for c in corners],
            [{tag(k): up(v) for k, v in s.items()} for s in sides],
            {tag(k): up(v) for k, v in op_center.items()})


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


def solve(corners, sides, op_center, x0, tol_residual=1e-10, tol_energy=1e-13, max_iter=400, check_every=4):
    """The plain FP64 form: returns x (N-normalised), the Ritz value, the energy (FP64 Rayleigh quotient),
    the residual estimate and the Lanczos iteration count."""
    H, N = S2.formExpectationAndNormalizationMultipliers(corners, sides, op_center)
    solver = GroundStateSolver(H.operator, N.operator)
    x, theta, resid, iters = solver.solve(x0, tol_residual=tol_residual, tol_energy=tol_energy,
                                          max_iter=max_iter, check_every=check_every)
    energy = solver.rayleigh_quotient(x)
    return dict(x=x, theta=theta, energy=energy, residual=resid, iterations=iters, operator=H.operator)


def solve_mixed(corners, sides, op_center, x0, dtype="f16", **kw):
    """The precision form of the same solve: stage 1/2 and the normalization factor in FP64, FP16/BF16
    Lanczos on the whitened operator, FP64 Jacobi-Davidson refinement.  The energy is the FP64
    Rayleigh-Ritz value of the refined vector (its Rayleigh quotient, from FP64 H and N products)."""
    H, N = S2.formExpectationAndNormalizationMultipliers(corners, sides, op_center)
    solver = MixedSolver(H.operator, N.operator, dtype=dtype)
    x, rho, info = solver.solve(x0, **kw)
    return dict(x=x, energy=rho, iterations=info["lowprec_iterations"], info=info, operator=H.operator,
                lp_operator=solver.Hlp)
