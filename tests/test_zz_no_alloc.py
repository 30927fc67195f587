"""The allocation-free stage-3 entry point (cc_stage3_matvec_fused_f64_noalloc) against the workspace
entry point it mirrors (cc_stage3_matvec_fused_f64), through the C ABI: the same kernels on the same
plan, so the results are bit-identical.  Two calls in a row on one device, the second on a smaller
operator (the scratch keeps the first call's partial sums, which the second must not see), then the
first operator again; a problem whose partial sums exceed the scratch is refused before any launch."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _operator(D, chi, model, seed):
    from carcassonne_amd import synthetic as syn
    from carcassonne_amd.tensors._2d import sparse as PS
    from tests.oracle_bridge import to_product
    c, s, st = syn.make_environment(D, chi, model=model, seed=seed)
    H, _ = PS.formExpectationAndNormalizationMultipliers(*to_product(c, s, syn.operator_center(model)))
    op = H.operator
    op._ensure_fused()
    return op, torch.from_numpy(st.ravel()).cuda()


def _noalloc(op, x, y, n_alpha=None):
    from carcassonne_amd import _lib
    from carcassonne_amd.device import ptr, stream
    f = op._fused
    _lib.call("cc_stage3_matvec_fused_f64_noalloc", stream(), f["n_a"], f["A_ptrs"], f["n_b"], f["B_ptrs"],
              f["n_terms"], f["ta"], f["tb"], f["to"], f["n_ops"], f["ops"],
              op.n_alpha if n_alpha is None else n_alpha, op.P, op.d, ptr(x), ptr(y))


def test_noalloc_matvec_equals_workspace_matvec(gpu):
    big, xb = _operator(6, 24, "heisenberg", 21)       # n_alpha = 576, 14 terms
    small, xs = _operator(4, 12, "ising", 22)          # n_alpha = 144, 7 terms
    for op, x in ((big, xb), (small, xs), (big, xb)):
        print("case D={} n_alpha={} terms={}".format(int(round(op.P ** 0.5)), op.n_alpha, len(op._A)), flush=True)
        want = torch.empty_like(x)
        op.apply_flat(x, want)
        got = torch.full_like(x, float("nan"))
        _noalloc(op, x, got)
        torch.cuda.synchronize()
        assert torch.equal(got, want)


def test_noalloc_rejects_sizes_beyond_the_scratch(gpu):
    op, x = _operator(6, 8, "ising", 23)
    y = torch.zeros_like(x)
    with pytest.raises(ValueError, match="scratch"):
        _noalloc(op, x, y, n_alpha=16 * 1024)          # 1024 groups x 2592 doubles > 1048576: refused on the host
    assert torch.count_nonzero(y) == 0
