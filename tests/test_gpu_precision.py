"""GPU tests of the precision form: FP16/BF16 stage-3 matvecs and the mixed-precision solve.
The low-precision matvec is checked against the FP64 fused kernel with a tolerance set by the
operand rounding (2^-8 BF16, 2^-11 FP16 relative per element, FP32 accumulation); the solve is
checked against the oracle's generalized eigh energy (tests/golden/energies.json) to 1e-10."""
import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "energies.json")))


def _operators(D, chi, model, seed):
    from carcassonne_amd import synthetic as syn
    from carcassonne_amd.tensors._2d import sparse as PS
    from tests.oracle_bridge import to_product
    c, s, st = syn.make_environment(D, chi, model=model, seed=seed)
    H, N = PS.formExpectationAndNormalizationMultipliers(*to_product(c, s, syn.operator_center(model)))
    return H.operator, N.operator, st


@pytest.mark.parametrize("split", [False, True, "vector"])
@pytest.mark.parametrize("dtype,tol", [("bf16", 2e-2), ("f16", 3e-3)])
@pytest.mark.parametrize("D,chi", [(2, 8), (4, 16), (5, 12), (6, 20)])
def test_lowprec_matvec_close_to_fp64(gpu, dtype, tol, D, chi, split):
    from carcassonne_amd.precision import LowPrecisionOperator
    H, N, st = _operators(D, chi, "ising", 3)
    x = torch.from_numpy(st.ravel()).cuda()
    for op in (H, N):
        y64, ylp = torch.empty_like(x), torch.empty_like(x)
        op.apply_flat(x, y64)
        LowPrecisionOperator(op, dtype, split=split).apply_flat(x, ylp)
        err = float((ylp - y64).norm() / y64.norm())
        assert err < tol, (dtype, D, err)
        assert err > 0.0   # really reduced precision


@pytest.mark.parametrize("D,chi", [(4, 16), (6, 12)])
def test_lowprec_matvec_any_alpha_grouping(gpu, D, chi):
    """The FP16 kernel double-buffers the alphas of each workgroup; workgroups of 1, 2, 3, 4, 5 and
    uneven alpha counts (and more workgroups than alphas) give the FP64 result to the FP16 tolerance and
    the default grouping's result up to FP32 summation order."""
    from carcassonne_amd.precision import LowPrecisionOperator
    H, N, st = _operators(D, chi, "ising", 5)
    x = torch.from_numpy(st.ravel()).cuda()
    y64, ydef = torch.empty_like(x), torch.empty_like(x)
    H.apply_flat(x, y64)
    base = LowPrecisionOperator(H, "f16")
    base.apply_flat(x, ydef)
    na = base.n_alpha
    for groups in (na, na // 2, na // 3, na // 4 + 1, na // 5, 7, 1, min(2 * na, 2048)):
        y = torch.empty_like(x)
        LowPrecisionOperator(H, "f16", groups=groups).apply_flat(x, y)
        assert float((y - y64).norm() / y64.norm()) < 3e-3, (D, groups)
        assert float((y - ydef).norm() / ydef.norm()) < 1e-5, (D, groups)


@pytest.mark.parametrize("dtype", ["bf16", "f16"])
@pytest.mark.parametrize("key", ["D4_chi32_ising_seed0", "D4_chi32_heisenberg_seed0"])
def test_mixed_solve_reaches_fp64_energy(gpu, dtype, key):
    from carcassonne_amd.precision import MixedSolver
    g = GOLDEN[key]
    H, N, st = _operators(g["D"], g["chi"], g["model"], g["seed"])
    solver = MixedSolver(H, N, dtype=dtype)
    x, rho, info = solver.solve(torch.from_numpy(st.ravel()).cuda())
    err = abs(rho - g["ground_energy"]) / abs(g["ground_energy"])
    print(dtype, key, "err", err, info)
    assert err < 1e-10, (dtype, key, rho, g["ground_energy"], err, info)


def test_sampled_launch_timing_counts(gpu):
    """cc_profile_begin_sampled(k) times launches 0, k, 2k, ... of each kind: 10 low-precision applies at
    k = 4 give 3 timed launches with a positive duration, and no FP64 launches."""
    from carcassonne_amd import _lib
    from carcassonne_amd.precision import LowPrecisionOperator
    H, N, st = _operators(4, 16, "ising", 7)
    x = torch.from_numpy(st.ravel()).cuda()
    y = torch.empty_like(x)
    lp = LowPrecisionOperator(H, "f16")
    lp.apply_flat(x, y)
    torch.cuda.synchronize()
    _lib.call("cc_profile_begin_sampled", 4)
    for _ in range(10):
        lp.apply_flat(x, y)
    kms, launches = _lib.f64_array([0, 0]), _lib.i64_array([0, 0])
    _lib.call("cc_profile_end", 2, kms, launches)
    assert list(launches) == [0, 3]
    assert kms[1] > 0.0
    with pytest.raises(Exception):
        _lib.call("cc_profile_begin_sampled", 0)


@pytest.mark.parametrize("model", ["ising", "heisenberg"])
def test_lowprec_plan_runs_and_chunks_at_d6(gpu, model):
    """The FP16 kernel reads its term and run tables by lane and loads the next term's and run's operands
    ahead (across runs and chunk boundaries): at D=6 (P=36), with the default grouping and with one
    workgroup per alpha pair, it gives the FP64 result to the FP16 tolerance.  TFIM (7 terms) fits one
    chunk; Heisenberg (14 terms) needs two, so the chunk-boundary path runs."""
    from carcassonne_amd.precision import LowPrecisionOperator
    H, N, st = _operators(6, 12, model, 9)
    x = torch.from_numpy(st.ravel()).cuda()
    y64 = torch.empty_like(x)
    H.apply_flat(x, y64)
    for groups in (None, H.n_alpha // 2):
        lp = LowPrecisionOperator(H, "f16", groups=groups)
        y = torch.empty_like(x)
        lp.apply_flat(x, y)
        err = float((y - y64).norm() / y64.norm())
        print(model, "terms", len(H._A), "chunks", lp.chunks, "groups", lp.groups, "err", err)
        assert lp.chunks == (1 if model == "ising" else 2), (model, lp.chunks)
        assert err < 3e-3, (model, groups, err)


@pytest.mark.parametrize("split", ["none", "vector", "both"])
@pytest.mark.parametrize("D,chi", [(4, 16), (6, 12)])
def test_lowprec_full_coefficient_terms(gpu, D, chi, split):
    """A site operator with a full 2x2 coefficient matrix (a longitudinal plus transverse field,
    O = -3 (X + 0.3 Z)) added to the Heisenberg model gives a plan that mixes diagonal (Z), anti-diagonal
    (X, iY) and full terms, so the low-precision kernel takes its general-coefficient path (over two term
    chunks at D=6): in every split mode it gives the FP64 fused result to the rounding tolerance."""
    from carcassonne_amd import synthetic as syn
    from carcassonne_amd.precision import LowPrecisionOperator
    from carcassonne_amd.tensors._2d import sparse as PS
    from tests.oracle_bridge import to_product
    c, s, st = syn.make_environment(D, chi, model="heisenberg", seed=13)
    op = syn.operator_center("heisenberg")
    op[("O",)] = -3.0 * (syn.X + 0.3 * syn.Z)
    H, _ = PS.formExpectationAndNormalizationMultipliers(*to_product(c, s, op))
    H = H.operator
    x = torch.from_numpy(st.ravel()).cuda()
    y64 = torch.empty_like(x)
    H.apply_flat(x, y64)
    lp = LowPrecisionOperator(H, "f16", split=split)
    y = torch.empty_like(x)
    lp.apply_flat(x, y)
    err = float((y - y64).norm() / y64.norm())
    print(D, chi, split, "terms", len(H._A), "chunks", lp.chunks, "diag form", lp.diagonal_form, "err", err)
    assert not lp.diagonal_form
    assert lp.chunks == (2 if D == 6 else 1)
    assert 0.0 < err < (3e-2 if split == "none" else 3e-3), (D, split, err)
