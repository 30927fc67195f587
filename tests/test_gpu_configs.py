"""GPU parity at the benchmark workloads' own sizes, against the oracle's committed golden values
(tests/golden/energies.json, tests/golden/batched_energies.json, written by tests/golden/make_golden.py):

* D=6, chi=64 (TFIM h=3.0, w=3, 7 terms; Heisenberg J=1, w=5, 14 terms): the FP64 fused matvec against
  the oracle's H.v checksums, and the mixed-precision solve (FP16 and BF16 matvecs, FP64 refinement)
  against the oracle's generalized-eigh ground energy to 1e-10 relative;
* D=4, chi=32: the plain FP64 Lanczos against the same energies, and relaxOver's control flow
  (restarts, multiplications, value) against the oracle's restatement;
* the batched workload: 64 systems over (D, chi) in {(3,24), (4,32), (5,48), (6,64)}, h in U[2, 4],
  BF16 matvecs, one HIP-graph launch, every energy against the oracle's to 1e-10.
"""
import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(__file__)
GOLDEN = json.load(open(os.path.join(HERE, "golden", "energies.json")))
BATCHED = json.load(open(os.path.join(HERE, "golden", "batched_energies.json")))


def _operators(key):
    from carcassonne_amd import synthetic as syn
    from carcassonne_amd.tensors._2d import sparse as PS
    from tests.oracle_bridge import to_product
    g = GOLDEN[key]
    c, s, st = syn.make_environment(g["D"], g["chi"], model=g["model"], seed=g["seed"])
    H, N = PS.formExpectationAndNormalizationMultipliers(*to_product(c, s, syn.operator_center(g["model"])))
    return g, H.operator, N.operator, st


@pytest.mark.parametrize("key", ["D6_chi64_ising_seed0", "D6_chi64_heisenberg_seed0"])
def test_fp64_matvec_at_d6_chi64_matches_oracle_checksums(gpu, key):
    """The FP64 fused stage-3 kernel on the seeded state: |Hv|, sum(Hv), the first 8 entries of Hv and
    |Nv| equal the oracle's to FP64 round-off."""
    g, H, N, st = _operators(key)
    print(key, "terms", len(H._A), flush=True)
    assert H.use_fused and len(H._A) == g["terms"]
    x = torch.from_numpy(st.ravel()).cuda()
    hv, nv = torch.empty_like(x), torch.empty_like(x)
    H.apply_flat(x, hv)
    N.apply_flat(x, nv)
    hv, nv = hv.cpu().numpy(), nv.cpu().numpy()
    assert abs(np.linalg.norm(hv) - g["Hv_norm"]) <= 1e-12 * g["Hv_norm"]
    assert abs(np.linalg.norm(nv) - g["Nv_norm"]) <= 1e-12 * g["Nv_norm"]
    assert abs(hv.sum() - g["Hv_sum"]) <= 1e-12 * g["Hv_norm"] * np.sqrt(hv.size)
    np.testing.assert_allclose(hv[:8], g["Hv_head"], rtol=0, atol=1e-12 * g["Hv_norm"])
    e = float(st.ravel() @ hv / (st.ravel() @ nv))
    assert abs(e - g["state_energy"]) <= 1e-12 * abs(g["state_energy"])


@pytest.mark.parametrize("dtype", ["f16", "bf16"])
@pytest.mark.parametrize("key", ["D6_chi64_ising_seed0", "D6_chi64_heisenberg_seed0"])
def test_mixed_solve_at_d6_chi64_reaches_oracle_energy(gpu, key, dtype):
    """The precision form at the headline size: FP16 or BF16 matvecs with FP32 accumulation, FP64
    Jacobi-Davidson refinement; the FP64 Rayleigh-Ritz energy is within 1e-10 of the oracle's.  BF16's
    8-bit significand makes the correction equation's operator ~8x less accurate than FP16's, so each
    FP64 pass gains less (about 2.5x in energy error against about 200x): it gets up to 24 passes."""
    from carcassonne_amd.precision import MixedSolver
    g, H, N, st = _operators(key)
    solver = MixedSolver(H, N, dtype=dtype)
    print(key, dtype, "chunks", solver.Hlp.chunks, flush=True)
    x, rho, info = solver.solve(torch.from_numpy(st.ravel()).cuda(), max_outer=8 if dtype == "f16" else 24)
    err = abs(rho - g["ground_energy"]) / abs(g["ground_energy"])
    print(key, dtype, "energy", rho, "oracle", g["ground_energy"], "rel err", err, info, flush=True)
    assert solver.Hlp.chunks == (1 if g["model"] == "ising" else 2)
    assert err < 1e-10, (key, dtype, rho, g["ground_energy"], err, info)
    assert info["converged"] and info["stopped_by"] != "max_outer", info


@pytest.mark.parametrize("key", ["D4_chi32_ising_seed0", "D4_chi32_heisenberg_seed0"])
def test_plain_fp64_lanczos_at_d4_chi32(gpu, key):
    """The plain FP64 GPU form at D=4, chi=32: Lanczos to a 1e-10 Ritz residual lands on the oracle's
    ground energy, and the FP64 Rayleigh quotient of its vector agrees."""
    from carcassonne_amd.solver import GroundStateSolver
    g, H, N, st = _operators(key)
    solver = GroundStateSolver(H, N)
    x, theta, resid, iters = solver.solve(torch.from_numpy(st.ravel()).cuda(), tol_residual=1e-10)
    rq = solver.rayleigh_quotient(x)
    print(key, "theta", theta, "rq", rq, "oracle", g["ground_energy"], "iters", iters, flush=True)
    assert resid <= 1e-10 * abs(theta) or iters < solver.n
    assert abs(theta - g["ground_energy"]) <= 1e-10 * abs(g["ground_energy"])
    assert abs(rq - g["ground_energy"]) <= 1e-10 * abs(g["ground_energy"])


def test_relax_over_control_flow_at_d4_chi32(gpu):
    """relaxOver at D=4, chi=32 (n = 512): the device restatement makes the oracle's decisions (dense H
    by the Multiplier cost rule, dense LU-class N), the same number of restarts and multiplications, and
    reaches the oracle's final value and vector."""
    from carcassonne_amd import synthetic as syn
    from carcassonne_amd.tensors._2d import sparse as PS
    from carcassonne_amd.solver import relax_over
    from carcassonne_amd.data import HipData
    from oracle import sparse2d, solver as OS
    from tests.oracle_bridge import to_oracle, to_product
    c, s, st = syn.make_environment(4, 32, model="ising", seed=0)
    op = syn.operator_center("ising")
    ref = sparse2d.form_multipliers(*to_oracle(c, s, op))
    use_dense = OS.is_cheaper_to_form_matrix(6, ref.cost_H, ref.cost_H_form, ref.n)
    ostats = {}
    vo = OS.relax_over(st.ravel(), lambda x: ref.apply_H(x).ravel(), lambda x: ref.apply_N(x).ravel(), max_mults=100,
                       tol=1e-7, H_matrix=ref.matrix_H() if use_dense else None, N_matrix=ref.matrix_N(),
                       use_dense_H=use_dense, use_lu_N=True, stats=ostats)
    H, N = PS.formExpectationAndNormalizationMultipliers(*to_product(c, s, op))
    assert H.isCheaperToFormMatrix(6) == use_dense
    pstats = {}
    vp = relax_over(HipData.fromArray(st), H, N, maximum_number_of_multiplications=100, stats=pstats)
    print("oracle", ostats, "device", pstats, flush=True)
    assert ostats["restarts"] > 1
    assert pstats["restarts"] == ostats["restarts"] and pstats["mults"] == ostats["mults"]
    assert abs(pstats["value"] - ostats["value"].real) <= 1e-9 * abs(ostats["value"])
    assert abs(abs(np.vdot(vp.toArray().ravel(), vo)) - 1.0) < 1e-6


@pytest.mark.parametrize("dtype", ["bf16", "f16"])
def test_batched_64_systems_match_oracle(gpu, dtype):
    """The batched workload: 64 seeded systems of four shapes, BF16 (as specified) or FP16 matvecs,
    autotuned per shape bucket, every eigensolve in one HIP graph with an in-device convergence record
    (energy of every FP64 pass, final residual).  Two replays with eager work between them give identical
    energies.  Every system the record passes is within 1e-10 of the oracle's energy; the guard's
    fallback (an eager FP16 mixed solve) brings every flagged system there too."""
    from carcassonne_amd.batched import BatchedSolve, draw_specs
    from carcassonne_amd.device import synchronize
    specs = draw_specs(64, seed=0)
    want = BATCHED["systems"]
    assert [(s[0], s[1], s[3]) for s in specs] == [(w["D"], w["chi"], w["seed"]) for w in want]
    b = BatchedSolve(specs, dtype=dtype)
    b.prepare()
    b.tune_work()
    print(dtype, "tuning", {str(k): v for k, v in b.tuning.items()}, flush=True)
    b.capture()
    first = None
    for rep in range(2):
        b.replay()
        got = b.energies()
        errs = [abs(g - w["ground_energy"]) / abs(w["ground_energy"]) for g, w in zip(got, want)]
        conv = b.convergence()
        flagged = [i for i, c in enumerate(conv) if not c[0]]
        print(dtype, "replay", rep, "max rel err", max(errs), "flagged", flagged, flush=True)
        for i, (e, c) in enumerate(zip(errs, conv)):
            print("  system", i, specs[i][:2], "err {:.2e} change {:.2e} residual {:.2e}".format(e, c[1], c[2]),
                  "ok" if c[0] else "FLAGGED")
        assert all(errs[i] < 1e-10 for i, c in enumerate(conv) if c[0]), "the record passed an unconverged system"
        if first is None:
            first = got
        assert got == first
        fixed = b.resolve_unconverged(conv)
        for i, rho in fixed.items():
            assert abs(rho - want[i]["ground_energy"]) < 1e-10 * abs(want[i]["ground_energy"]), (i, rho)
        if dtype == "f16":
            assert not flagged
        synchronize()


@pytest.mark.parametrize("key", ["D6_chi64_ising_seed0", "D6_chi64_heisenberg_seed0"])
def test_operator_properties_at_d6_chi64(gpu, key):
    """Size-independent properties of the stage-3 operators at the headline size, for every kernel: H and
    N are linear (H(ax + by) = a Hx + b Hy) and symmetric (y.Hx = x.Hy; the environment is ket/bra
    symmetric), N is positive on random vectors, and the FP32 and FP16 kernels agree with the FP64 one to
    their rounding (FP32 operands 2e-6, FP16 slabs 3e-3 relative)."""
    from carcassonne_amd import _lib
    from carcassonne_amd.device import ptr, stream
    from carcassonne_amd.precision import LowPrecisionOperator
    g, H, N, st = _operators(key)
    rng = np.random.default_rng(7)
    x = torch.from_numpy(rng.standard_normal(H.n)).cuda()
    y = torch.from_numpy(rng.standard_normal(H.n)).cuda()
    a, b = 0.37, -1.9
    lp = LowPrecisionOperator(H, "f16")

    def f64(op):
        return lambda v: (lambda out: (op.apply_flat(v, out), out)[1])(torch.empty_like(v))

    def f32(op):
        h = op.plan_handle("f32")
        return lambda v: (lambda out: (_lib.call("cc_stage3_plan_apply", stream(), h, ptr(v), ptr(out)), out)[1])(torch.empty_like(v))

    def f16(v):
        out = torch.empty_like(v)
        lp.apply_flat(v, out)
        return out

    for name, apply, tol_lin in (("H f64", f64(H), 1e-13), ("N f64", f64(N), 1e-13), ("H f32", f32(H), 1e-6)):
        hx, hy, hxy = apply(x), apply(y), apply(a * x + b * y)
        lin = float((hxy - (a * hx + b * hy)).norm() / hxy.norm())
        sym = abs(float(y @ hx) - float(x @ hy)) / (float(hx.norm()) * float(y.norm()))
        print(key, name, "linearity %.1e symmetry %.1e" % (lin, sym), flush=True)
        assert lin < tol_lin and sym < 10 * tol_lin, (name, lin, sym)
    nx = f64(N)(x)
    assert float(x @ nx) > 0 and float(y @ f64(N)(y)) > 0
    h64 = f64(H)(x)
    assert float((f32(H)(x) - h64).norm() / h64.norm()) < 2e-6
    assert float((f16(x) - h64).norm() / h64.norm()) < 3e-3
