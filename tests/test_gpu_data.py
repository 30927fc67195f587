"""GPU parity of the device Data protocol (HipData) against NumPy / the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rng(seed=0):
    return np.random.default_rng(seed)


def test_contract_with_matches_tensordot(gpu):
    from carcassonne_amd.data import HipData
    rng = _rng(1)
    cases = [((3, 4, 5), (5, 4, 2), (1, 2), (1, 0)), ((6, 7), (7, 3), (1,), (0,)), ((2, 3, 4, 5), (4, 2, 6), (0, 2), (1, 0)),
             ((5, 3), (2, 4), (), ()), ((4, 4, 4), (4, 4, 4), (0, 1, 2), (2, 1, 0)), ((70, 65), (65, 130), (1,), (0,)),
             ((3, 1, 2, 5), (5, 3), (3, 0), (0, 1))]
    for sa, sb, aa, ab in cases:
        a, b = rng.standard_normal(sa), rng.standard_normal(sb)
        got = HipData.fromArray(a).contractWith(HipData.fromArray(b), aa, ab).toArray()
        want = np.tensordot(a, b, (aa, ab))
        assert got.shape == want.shape
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_gemm_tile_shapes_match_numpy(gpu):
    """cc_gemm_f64 on its tile shapes, 32 x 32 (small grids) and 64 x 64 (1296 x 1296 x 1100, the shape of
    the normalization environment's product; 1300 x 1290 leaves ragged edges on both tile dimensions),
    with K-contiguous and row-contiguous operands, alpha and beta."""
    import torch
    from carcassonne_amd import _lib
    from carcassonne_amd.device import ptr, stream
    rng = _rng(4)
    for M, N, K, a_t, b_t in ((40, 30, 50, False, False), (700, 600, 96, True, False),
                              (1296, 1296, 1100, True, False), (1300, 1290, 1030, False, True)):
        a = rng.standard_normal((K, M) if a_t else (M, K))
        b = rng.standard_normal((N, K) if b_t else (K, N))
        c0 = rng.standard_normal((M, N))
        A, B, C = (torch.from_numpy(x.copy()).cuda() for x in (a, b, c0))
        sAm, sAk = (1, M) if a_t else (K, 1)
        sBk, sBn = (1, K) if b_t else (N, 1)
        _lib.call("cc_gemm_f64", stream(), 1, M, N, K, 0.5, ptr(A), 0, sAm, sAk, ptr(B), 0, sBk, sBn, 2.0,
                  ptr(C), 0, N, 1)
        want = 0.5 * ((a.T if a_t else a) @ (b.T if b_t else b)) + 2.0 * c0
        np.testing.assert_allclose(C.cpu().numpy(), want, rtol=1e-12, atol=1e-11 * np.sqrt(K))


def test_gemm_batched_match_numpy(gpu):
    """cc_gemm_f64 batched with strides between the batch entries (a 1296 x 1296 product's sum split
    over four ranges of 1024 is 4 x 1296 x 1296 x 1024), K-contiguous and row-contiguous operands,
    ragged edges, alpha and beta."""
    import torch
    from carcassonne_amd import _lib
    from carcassonne_amd.device import ptr, stream
    rng = _rng(5)
    # the last two: single products of one to two 64 x 64 tiles per CU, the second with ragged edges
    for batch, M, N, K, a_t, b_t in ((4, 1296, 1296, 1024, True, False), (3, 1300, 1290, 600, False, True),
                                     (1, 1296, 1296, 700, True, False), (1, 1100, 1000, 300, False, True)):
        a = rng.standard_normal((batch, K, M) if a_t else (batch, M, K))
        b = rng.standard_normal((batch, N, K) if b_t else (batch, K, N))
        c0 = rng.standard_normal((batch, M, N))
        A, B, C = (torch.from_numpy(x.copy()).cuda() for x in (a, b, c0))
        sAm, sAk = (1, M) if a_t else (K, 1)
        sBk, sBn = (1, K) if b_t else (N, 1)
        _lib.call("cc_gemm_f64", stream(), batch, M, N, K, -0.75, ptr(A), M * K, sAm, sAk, ptr(B), K * N, sBk, sBn,
                  0.5, ptr(C), M * N, N, 1)
        at = a.transpose(0, 2, 1) if a_t else a
        bt = b.transpose(0, 2, 1) if b_t else b
        want = -0.75 * np.matmul(at, bt) + 0.5 * c0
        np.testing.assert_allclose(C.cpu().numpy(), want, rtol=1e-12, atol=1e-11 * np.sqrt(K))


def test_permute_awkward_extents_match_numpy(gpu):
    """cc_permute (the 32-bit coordinate path, extents divided by multiply-high) on rank-6 shapes with
    extents of 1, primes and large odd values, and on a strided source (every other element of a base)."""
    import torch
    from carcassonne_amd import _lib
    from carcassonne_amd.device import ptr, stream
    rng = _rng(6)
    for shape, perm in (((1, 7, 13, 3, 1, 5), (5, 2, 0, 4, 1, 3)), ((97, 1, 3, 11, 2, 1), (3, 0, 5, 2, 4, 1)),
                        ((4099, 3, 5), (2, 0, 1)), ((2, 3, 4, 5, 6, 7), (5, 4, 3, 2, 1, 0))):
        a = rng.standard_normal(shape)
        st = [int(np.prod(shape[k + 1:])) for k in range(len(shape))]
        out_shape = [shape[p] for p in perm]
        src = torch.from_numpy(a.copy()).cuda()
        dst = torch.empty(int(np.prod(out_shape)), dtype=torch.float64, device="cuda")
        _lib.call("cc_permute", stream(), _lib.CC_F64, ptr(src), ptr(dst), len(shape), _lib.i64_array(out_shape),
                  _lib.i64_array([st[p] for p in perm]))
        np.testing.assert_array_equal(dst.cpu().numpy().reshape(out_shape), a.transpose(perm))
    # strided source: the even elements of a (6, 10) base read as a (6, 5) matrix, written transposed
    base = rng.standard_normal((6, 10))
    src = torch.from_numpy(base.copy()).cuda()
    dst = torch.empty(30, dtype=torch.float64, device="cuda")
    _lib.call("cc_permute", stream(), _lib.CC_F64, ptr(src), ptr(dst), 2, _lib.i64_array([5, 6]), _lib.i64_array([2, 10]))
    np.testing.assert_array_equal(dst.cpu().numpy().reshape(5, 6), base[:, ::2].T)


def test_join_transpose_and_elementwise(gpu):
    from carcassonne_amd.data import HipData
    rng = _rng(2)
    a = rng.standard_normal((2, 3, 4, 5))
    h = HipData.fromArray(a)
    np.testing.assert_array_equal(h.join((0, 2), 3, 1).toArray(), a.transpose(0, 2, 3, 1).reshape(8, 5, 3))
    np.testing.assert_array_equal(h.transpose(3, 1, 0, 2).toArray(), a.transpose(3, 1, 0, 2))
    b = rng.standard_normal(a.shape)
    hb = HipData.fromArray(b)
    np.testing.assert_allclose((h + hb).toArray(), a + b, rtol=0, atol=0)
    np.testing.assert_allclose((h * hb).toArray(), a * b, rtol=0, atol=0)
    np.testing.assert_allclose((h * 2.5).toArray(), a * 2.5, rtol=0, atol=0)
    h2 = h.copy()
    h2 += hb
    np.testing.assert_allclose(h2.toArray(), a + b, rtol=0, atol=0)
    assert abs(h.norm() - np.linalg.norm(a)) <= 1e-12 * np.linalg.norm(a)
    assert not h.hasNaN()


def test_svd_and_normalize_axis(gpu):
    from carcassonne_amd.data import HipData
    from oracle import linalg as OL
    rng = _rng(3)
    for m, n in [(20, 6), (432, 6), (64, 64), (7, 1), (300, 37)]:
        a = rng.standard_normal((m, n))
        U, S, Vt = HipData.fromArray(a).svd(False)
        U, S, Vt = U.toArray(), S.toArray(), Vt.toArray()
        np.testing.assert_allclose(S, np.linalg.svd(a, compute_uv=False), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose((U * S) @ Vt, a, rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(U.T @ U, np.eye(n), atol=1e-12)
    # normalizeAxis: isometry, normalizer and denormalizer are gauge-invariant products
    s = rng.standard_normal((3, 2, 3, 2, 2))
    for axis in range(4):
        iso, nrm, den = HipData.fromArray(s).normalizeAxis(axis)
        oiso, onrm, oden = OL.normalize_axis(s, axis)
        np.testing.assert_allclose(den.toArray(), oden, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(nrm.toArray(), onrm, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(iso.toArray(), oiso, rtol=1e-10, atol=1e-12)


def test_svd_rank_deficient_has_orthonormal_u(gpu):
    """Zero singular values: U is completed to orthonormal columns (numpy's thin U is), so polar factors
    stay isometries; the product U S Vt still reproduces A."""
    from carcassonne_amd.data import HipData
    from oracle import linalg as OL
    rng = _rng(4)
    a1 = np.zeros((9, 5))
    a1[:, 1] = rng.standard_normal(9)            # rank 1 with four exactly-zero columns
    a2 = np.outer(rng.standard_normal(12), rng.standard_normal(6))   # rank 1, rounding-level columns
    a3 = np.zeros((6, 6))                        # all zero
    a4 = np.zeros((40, 8))
    a4[:3, :3] = rng.standard_normal((3, 3))     # block with zero rows and columns
    a5 = np.zeros((4, 4))
    a5[0, :] = 1.0                               # row of ones: e_0 is the covered row
    for a in (a1, a2, a3, a4, a5):
        m, n = a.shape
        U, S, Vt = HipData.fromArray(a).svd(False)
        U, S, Vt = U.toArray(), S.toArray(), Vt.toArray()
        np.testing.assert_allclose(S, np.linalg.svd(a, compute_uv=False), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(U.T @ U, np.eye(n), atol=1e-12)
        np.testing.assert_allclose(Vt @ Vt.T, np.eye(n), atol=1e-12)
        np.testing.assert_allclose((U * S) @ Vt, a, rtol=1e-11, atol=1e-11)
        Q = HipData.fromArray(a).unitize().toArray()
        np.testing.assert_allclose(Q.T @ Q, np.eye(n), atol=1e-12)
    # normalizeAxis of a tensor whose bond is rank deficient: the isometry is one, the gauge-invariant
    # normalizer and denormalizer match the oracle
    s = np.zeros((2, 3, 2, 4))
    s[..., 0] = rng.standard_normal((2, 3, 2))
    s[..., 2] = 2.0 * s[..., 0]
    iso, nrm, den = HipData.fromArray(s).normalizeAxis(3)
    iso = iso.toArray().reshape(-1, 4)
    np.testing.assert_allclose(iso.T @ iso, np.eye(4), atol=1e-12)
    _, onrm, oden = OL.normalize_axis(s, 3)
    np.testing.assert_allclose(den.toArray(), oden, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(nrm.toArray(), onrm, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("D,chi,model", [(2, 8, "ising"), (3, 6, "heisenberg"), (4, 8, "ising")])
def test_stage3_operator_matches_oracle(gpu, D, chi, model):
    from carcassonne_amd import synthetic as syn
    from carcassonne_amd.tensors._2d import sparse as PS
    from carcassonne_amd.data import HipData
    from oracle import sparse2d
    from tests.oracle_bridge import to_oracle, to_product, rel_err
    c, s, st = syn.make_environment(D, chi, model=model, seed=5)
    op = syn.operator_center(model)
    oc, os_, oo = to_oracle(c, s, op)
    ref = sparse2d.form_multipliers(oc, os_, oo)
    pc, ps, po = to_product(c, s, op)
    H, N = PS.formExpectationAndNormalizationMultipliers(pc, ps, po)
    v = HipData.fromArray(st)
    assert rel_err(H(v).toArray(), ref.apply_H(st)) < 1e-12
    assert rel_err(N(v).toArray(), ref.apply_N(st)) < 1e-12
    assert H.cost_of_multiply == ref.cost_H and N.cost_of_multiply == ref.cost_N
    assert H.cost_of_formMatrix == ref.cost_H_form
    if D <= 3:
        assert rel_err(H.formMatrix().toArray(), ref.matrix_H()) < 1e-12
        assert rel_err(N.formMatrix().toArray(), ref.matrix_N()) < 1e-12


@pytest.mark.parametrize("D,chi,model", [(2, 6, "ising"), (3, 5, "heisenberg"), (4, 6, "ising"), (5, 4, "ising"),
                                         (6, 5, "heisenberg"), (6, 9, "ising")])
def test_fused_matvec_matches_generic_and_oracle(gpu, D, chi, model):
    """The fused MFMA kernel, the generic GEMM path and the oracle agree on every instantiated shape."""
    import torch
    from carcassonne_amd import synthetic as syn
    from carcassonne_amd.tensors._2d import sparse as PS
    from carcassonne_amd.data import HipData
    from oracle import sparse2d
    from tests.oracle_bridge import to_oracle, to_product, rel_err
    c, s, st = syn.make_environment(D, chi, model=model, seed=11)
    op = syn.operator_center(model)
    ref = sparse2d.form_multipliers(*to_oracle(c, s, op))
    H, N = PS.formExpectationAndNormalizationMultipliers(*to_product(c, s, op))
    for M, apply_ref in ((H, ref.apply_H), (N, ref.apply_N)):
        o = M.operator
        assert o.use_fused
        x = HipData.fromArray(st)._t.reshape(-1)
        y1 = torch.empty_like(x)
        y2 = torch.empty_like(x)
        o.apply_flat(x, y1)
        o.apply_generic_flat(x, y2)
        want = apply_ref(st).ravel()
        assert rel_err(y1.cpu().numpy(), want) < 1e-12
        assert rel_err(y2.cpu().numpy(), want) < 1e-12


def test_strided_slab_packing_matches_join(gpu):
    """cc_stage3_pack_strided_f64 / cc_stage3_lp_convert_strided read a stage-2 tensor (contiguous, or a
    lazily permuted view of a contraction result) in place: same packed operand order as joining it to
    (alpha, P, P) slabs (stage2_0.join((0,1),4,5,2,3), stage2_1.join((1,0),4,5,2,3)) and packing, bit for
    bit, and the same rounded FP16 slabs and scale as the contiguous conversion."""
    import torch
    from carcassonne_amd import _lib
    from carcassonne_amd.data import HipData, PermutedView
    from carcassonne_amd.device import ptr, stream, empty
    from carcassonne_amd.stage3 import _Slab
    rng = _rng(21)
    chi, D = 5, 3
    P = D * D
    NT, KS = (P + 15) // 16, (P + 3) // 4
    raw = rng.standard_normal((chi, D, D, chi, D, D))            # legs (P1, P2, P3, Q0, Q2, Q3)
    perm = [3, 0, 1, 4, 2, 5]                                    # -> (Q0, P1, P2, Q2, P3, Q3)
    t_np = raw.transpose(perm)
    for t in (HipData.fromArray(np.ascontiguousarray(t_np)), PermutedView(torch.from_numpy(raw).cuda(), perm)):
        for which in (0, 1):
            joined = (t_np.transpose(0, 1, 4, 5, 2, 3) if which == 0 else t_np.transpose(1, 0, 4, 5, 2, 3)).reshape(chi * chi, P, P)
            want = np.zeros((chi * chi, KS, NT, 64))
            for u in range(KS):
                for n in range(NT):
                    for lane in range(64):
                        row, col = 16 * n + lane % 16, 4 * u + lane // 16
                        if row < P and col < P:
                            want[:, u, n, lane] = joined[:, row, col]
            sl = _Slab(t, which)
            np.testing.assert_array_equal(sl.packed().cpu().numpy().reshape(want.shape), want)
            np.testing.assert_array_equal(sl.joined.toArray().reshape(joined.shape), joined)
            # low-precision conversion: strided read equals the contiguous one
            n_el = _lib.query("cc_stage3_lp_slab_size", chi * chi, P)
            got, ref = torch.empty(n_el, dtype=torch.float16, device="cuda"), torch.empty(n_el, dtype=torch.float16, device="cuda")
            sc1, sc2 = empty((513,)), empty((513,))
            base, (na0, na1, d0, d1), st = sl.source()
            jd = torch.from_numpy(np.ascontiguousarray(joined)).cuda()
            _lib.call("cc_stage3_lp_convert", stream(), _lib.CC_F16, chi * chi, P, ptr(jd), ptr(sc2), ptr(ref))
            for src_elems in (base.numel(), 0):     # contiguous maximum, and the maximum through the map
                got.zero_()
                _lib.call("cc_stage3_lp_convert_strided", stream(), _lib.CC_F16, na0, na1, d0, d1, _lib.i64_array(st),
                          src_elems, ptr(base), ptr(sc1), ptr(got))
                assert float(sc1[0]) == float(sc2[0]) == float(np.abs(joined).max())
                assert torch.equal(got, ref)
