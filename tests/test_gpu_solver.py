"""GPU parity of the solve: dense linear algebra, the device Lanczos and the reference-compatible relaxOver."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_cholesky_trtri_skinny(gpu):
    from carcassonne_amd import _lib
    from carcassonne_amd.device import ptr, stream, empty
    rng = np.random.default_rng(7)
    for n in (5, 64, 65, 200, 333, 1296):
        g = rng.standard_normal((n, n))
        a = g @ g.T + n * np.eye(n)
        A = torch.from_numpy(a.copy()).cuda()
        info = torch.zeros(1, dtype=torch.int32, device=A.device)
        work = empty((_lib.query("cc_linalg_work_size", n),))
        _lib.call("cc_cholesky_f64", stream(), n, ptr(A), ptr(info), ptr(work))
        assert int(info.item()) == 0
        L = A.cpu().numpy()
        np.testing.assert_allclose(L, np.linalg.cholesky(a), rtol=1e-11, atol=1e-11)
        X = empty((n, n))
        _lib.call("cc_trtri_lower_f64", stream(), n, ptr(A), ptr(X), ptr(work))
        np.testing.assert_allclose(X.cpu().numpy() @ L, np.eye(n), atol=1e-11)
        Ad = torch.from_numpy(a).cuda()      # keep device copies referenced until the kernels ran
        for tri, m in ((0, a), (1, np.tril(a)), (2, np.triu(a))):
            for nr in (1, 2, 3):
                x = rng.standard_normal((n, nr))
                xd = torch.from_numpy(x).cuda()
                Y = empty((n, nr))
                _lib.call("cc_skinny_f64", stream(), n, tri, ptr(Ad), nr, ptr(xd), ptr(Y))
                np.testing.assert_allclose(Y.cpu().numpy(), m @ x, rtol=1e-12, atol=1e-10)
    # a non positive definite matrix is reported, not factored silently
    A = torch.from_numpy(-np.eye(8)).cuda()
    info = torch.zeros(1, dtype=torch.int32, device=A.device)
    work = empty((_lib.query("cc_linalg_work_size", 8),))
    _lib.call("cc_cholesky_f64", stream(), 8, ptr(A), ptr(info), ptr(work))
    assert int(info.item()) == 1


def test_cholesky_inverse_recursive(gpu):
    """cc_cholesky_inverse_f64: L equals numpy's Cholesky factor, X = L^-1 (X L = I), both lower with a
    zero upper triangle, for sizes that are and are not multiples of the 64-row leaves; a non-positive
    pivot inside a later leaf is reported with its 1-based global index by both Cholesky entry points."""
    from carcassonne_amd import _lib
    from carcassonne_amd.device import ptr, stream, empty
    rng = np.random.default_rng(11)
    for n in (1, 7, 64, 65, 128, 191, 333, 1296):
        g = rng.standard_normal((n, n))
        a = g @ g.T + n * np.eye(n)
        A = torch.from_numpy(a.copy()).cuda()
        X = empty((n, n))
        info = torch.full((1,), 5, dtype=torch.int32, device=A.device)
        work = empty((_lib.query("cc_linalg_work_size", n),))
        _lib.call("cc_cholesky_inverse_f64", stream(), n, ptr(A), ptr(X), ptr(info), ptr(work))
        assert int(info.item()) == 0
        L, Xh = A.cpu().numpy(), X.cpu().numpy()
        want = np.linalg.cholesky(a)
        np.testing.assert_allclose(L, want, rtol=1e-11, atol=1e-11)
        assert np.all(np.triu(L, 1) == 0.0) and np.all(np.triu(Xh, 1) == 0.0)
        np.testing.assert_allclose(Xh @ want, np.eye(n), atol=1e-11)
        # the recursive trtri on the same factor agrees
        X2 = empty((n, n))
        _lib.call("cc_trtri_lower_f64", stream(), n, ptr(A), ptr(X2), ptr(work))
        np.testing.assert_allclose(X2.cpu().numpy(), Xh, rtol=1e-12, atol=1e-12)
    for n, bad in ((200, 131), (333, 64), (100, 100)):
        a = np.eye(n)
        a[bad - 1, bad - 1] = -1.0
        for name in ("cc_cholesky_inverse_f64", "cc_cholesky_f64"):
            A = torch.from_numpy(a.copy()).cuda()
            info = torch.zeros(1, dtype=torch.int32, device=A.device)
            work = empty((_lib.query("cc_linalg_work_size", n),))
            if name == "cc_cholesky_f64":
                _lib.call(name, stream(), n, ptr(A), ptr(info), ptr(work))
            else:
                X = empty((n, n))
                _lib.call(name, stream(), n, ptr(A), ptr(X), ptr(info), ptr(work))
            assert int(info.item()) == bad, (name, n, bad, int(info.item()))


@pytest.mark.parametrize("D,chi,model,seed", [(2, 8, "ising", 0), (3, 6, "heisenberg", 1), (4, 8, "ising", 2),
                                              (4, 16, "heisenberg", 3)])
def test_lanczos_matches_generalized_eigh(gpu, D, chi, model, seed):
    """Pinned against scipy.linalg.eigh(H, N) -- minimizeExpectationUsingFullEigensolver (system/_2d.py:498-502)."""
    from scipy.linalg import eigh
    from carcassonne_amd import synthetic as syn
    from carcassonne_amd.tensors._2d import sparse as PS
    from carcassonne_amd.solver import GroundStateSolver
    from carcassonne_amd.data import HipData
    from oracle import sparse2d
    from tests.oracle_bridge import to_oracle, to_product
    c, s, st = syn.make_environment(D, chi, model=model, seed=seed)
    op = syn.operator_center(model)
    ref = sparse2d.form_multipliers(*to_oracle(c, s, op))
    w, V = eigh(ref.matrix_H(), ref.matrix_N())
    H, N = PS.formExpectationAndNormalizationMultipliers(*to_product(c, s, op))
    solver = GroundStateSolver(H.operator, N.operator)
    x, theta, resid, iters = solver.solve(HipData.fromArray(st), tol_residual=1e-11, tol_energy=0.0, max_iter=600)
    assert abs(theta - w[0]) <= 1e-10 * abs(w[0]), (theta, w[0], iters)
    rq = solver.rayleigh_quotient(x)
    assert abs(rq - w[0]) <= 1e-10 * abs(w[0])
    xv = x.cpu().numpy()
    Hm, Nm = ref.matrix_H(), ref.matrix_N()
    assert abs(xv @ Nm @ xv - 1.0) < 1e-10          # N-normalised
    r = Hm @ xv - theta * (Nm @ xv)
    assert np.linalg.norm(r) <= 1e-8 * abs(theta) * np.linalg.norm(Nm @ xv)


@pytest.mark.parametrize("D,chi,model,seed", [(2, 8, "ising", 3), (3, 5, "heisenberg", 4)])
def test_relax_over_control_flow_matches_oracle(gpu, D, chi, model, seed):
    """relaxOver on device vectors takes the same decisions as the oracle's restatement
    (dense-vs-matrix-free, restarts, multiplications) and lands on the same value."""
    from carcassonne_amd import synthetic as syn
    from carcassonne_amd.tensors._2d import sparse as PS
    from carcassonne_amd.solver import relax_over
    from carcassonne_amd.data import HipData
    from oracle import sparse2d, solver as OS
    from tests.oracle_bridge import to_oracle, to_product
    c, s, st = syn.make_environment(D, chi, model=model, seed=seed)
    op = syn.operator_center(model)
    ref = sparse2d.form_multipliers(*to_oracle(c, s, op))
    use_lu = OS.is_cheaper_to_form_matrix(60, ref.cost_N, ref.cost_N_form, ref.n)
    use_dense = OS.is_cheaper_to_form_matrix(6, ref.cost_H, ref.cost_H_form, ref.n)
    ostats = {}
    vo = OS.relax_over(st.ravel(), lambda x: ref.apply_H(x).ravel(), lambda x: ref.apply_N(x).ravel(), max_mults=100,
                       tol=1e-7, H_matrix=ref.matrix_H() if use_dense else None, N_matrix=ref.matrix_N(),
                       use_dense_H=use_dense, use_lu_N=True, stats=ostats)
    H, N = PS.formExpectationAndNormalizationMultipliers(*to_product(c, s, op))
    assert H.isCheaperToFormMatrix(6) == use_dense
    pstats = {}
    vp = relax_over(HipData.fromArray(st), H, N, maximum_number_of_multiplications=100, stats=pstats)
    assert pstats["restarts"] == ostats["restarts"] and pstats["mults"] == ostats["mults"]
    assert abs(pstats["value"] - ostats["value"].real) <= 1e-9 * abs(ostats["value"])
    assert abs(abs(np.vdot(vp.toArray().ravel(), vo)) - 1.0) < 1e-6


def test_non_positive_normalization_is_reported_by_the_solve(gpu):
    """The Cholesky pivot check of N_env is deferred to the solver's own synchronization; a negative
    definite normalization operator still ends the solve with ValueError (not a silent energy)."""
    from carcassonne_amd import synthetic as syn, stage3
    from carcassonne_amd.tensors._2d import sparse as PS
    from carcassonne_amd.solver import GroundStateSolver
    from tests.oracle_bridge import to_product
    c, s, st = syn.make_environment(2, 6, model="ising", seed=4)
    H, N = PS.formExpectationAndNormalizationMultipliers(*to_product(c, s, syn.operator_center("ising")))
    s0, s1, _ = N.operator.terms[0]
    negated = stage3.Stage3Operator([(s0 * -1.0, s1, None)], physical_dimension=N.operator.d)
    with pytest.raises(ValueError, match="not positive definite"):
        GroundStateSolver(H.operator, negated).solve(torch.from_numpy(st.ravel()).cuda())


def test_stage12_graph_replays_match_eager(gpu):
    """The replayed stage-1/2 graph gives the eager stage 1/2 on every replay: two environments of one
    structure alternate (A, B, A, B) with eager work between the replays, and each replay's H and N
    products on a seeded vector, and its mixed-precision energy, equal the eager path's."""
    from carcassonne_amd import pipeline, synthetic as syn
    from carcassonne_amd.data import HipData
    envs = []
    for seed in (3, 4):
        c, s, st = syn.make_environment(4, 16, model="ising", seed=seed)
        envs.append((pipeline.device_environment(c, s, syn.operator_center("ising", h=2.5 + seed / 10)), st))
    pipeline._STAGE12.clear()
    rng = np.random.default_rng(0)
    for rep, i in enumerate((0, 1, 0, 1)):
        (corners, sides, opc), st = envs[i]
        Hg, Ng = pipeline.multipliers(corners, sides, opc, graph=True)
        # eager work between replays, then the eager reference
        He, Ne = pipeline.multipliers(corners, sides, opc, graph=False)
        torch.cuda.synchronize()
        x = torch.from_numpy(rng.standard_normal(Hg.operator.n)).cuda()
        for g, e in ((Hg.operator, He.operator), (Ng.operator, Ne.operator)):
            yg, ye = torch.empty_like(x), torch.empty_like(x)
            g.apply_flat(x, yg)
            e.apply_flat(x, ye)
            assert torch.allclose(yg, ye, rtol=1e-12, atol=1e-12 * float(ye.abs().max())), (rep, i)
        x0 = HipData.fromArray(st)
        rg = pipeline.solve_mixed(corners, sides, opc, x0, graph=True)
        re = pipeline.solve_mixed(corners, sides, opc, x0, graph=False)
        assert abs(rg["energy"] - re["energy"]) <= 1e-11 * abs(re["energy"]), (rep, i, rg["energy"], re["energy"])
    assert len(pipeline._STAGE12) == 1
