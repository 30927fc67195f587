"""The FP32 form of the plain method: the fused stage-3 matvec with FP32 operands on v_mfma_f32_16x16x4f32
(FP32 accumulation, workgroup partials summed in FP64) against the FP64 fused kernel, and the Lanczos solve
on it against the oracle's ground energies at the plain GPU workload (D=4, chi=32) and the headline size
(D=6, chi=64).  The matvec tolerance is set by FP32 rounding: each slab entry and vector entry carries
2^-24 relative error, and the sums over alpha (chi^2 slabs) and p, q (P = D^2) accumulate in FP32."""
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


@pytest.mark.parametrize("model", ["ising", "heisenberg"])
@pytest.mark.parametrize("D,chi", [(2, 8), (3, 6), (4, 16), (5, 12), (6, 20)])
def test_fp32_matvec_close_to_fp64(gpu, D, chi, model):
    from carcassonne_amd import _lib
    from carcassonne_amd.device import ptr, stream
    H, N, st = _operators(D, chi, model, 3)
    rng = np.random.default_rng(D)
    for x_host in (st.ravel(), rng.standard_normal(H.n)):
        x = torch.from_numpy(np.ascontiguousarray(x_host)).cuda()
        for op in (H, N):
            y64, y32 = torch.empty_like(x), torch.empty_like(x)
            op.apply_flat(x, y64)
            _lib.call("cc_stage3_plan_apply", stream(), op.plan_handle("f32"), ptr(x), ptr(y32))
            err = float((y32 - y64).norm() / y64.norm())
            assert 0.0 < err < 2e-6, (D, chi, model, err)


@pytest.mark.parametrize("key", ["D4_chi32_ising_seed0", "D4_chi32_heisenberg_seed0", "D6_chi64_ising_seed0"])
def test_fp32_lanczos_energy(gpu, key):
    """Lanczos on the FP32 operator to its own 1e-10 Ritz residual lands within 1e-6 of the oracle's energy
    (the FP32 operator's eigenvalue); the FP64 Rayleigh quotient of its vector is accurate to second
    order in the vector's error, within 1e-9."""
    from carcassonne_amd.solver import GroundStateSolver
    g = GOLDEN[key]
    H, N, st = _operators(g["D"], g["chi"], g["model"], g["seed"])
    solver = GroundStateSolver(H, N, precision="f32")
    x, theta, resid, iters = solver.solve(torch.from_numpy(st.ravel()).cuda(), tol_residual=1e-10)
    rq = solver.rayleigh_quotient(x)
    e = g["ground_energy"]
    print(key, "theta", theta, "rq", rq, "oracle", e, "theta err %.2e rq err %.2e" % (abs(theta - e) / abs(e), abs(rq - e) / abs(e)),
          "iters", iters, flush=True)
    assert abs(theta - e) <= 1e-6 * abs(e)
    assert abs(rq - e) <= 1e-9 * abs(e)
