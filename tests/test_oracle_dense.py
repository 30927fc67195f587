"""Every dense index map of the oracle against an explicit-loop evaluation of the reference's join
spec on tiny random shapes (the reference's own strategy, tests/test_tensors_dense.py:10-187).
The loops name each output element's sum directly, so a swapped leg, a transposed operand or a
missed conjugate fails here."""
import itertools

import numpy as np
import pytest

from oracle import dense
from oracle.tags import L, R, O

RNG = np.random.default_rng(11)


def rnd(*shape, cplx=True):
    a = RNG.standard_normal(shape)
    return a + 1j * RNG.standard_normal(shape) if cplx else a


def loop(out_shape, sum_shape, f):
    out = np.zeros(out_shape, dtype=complex)
    for o in itertools.product(*map(range, out_shape)):
        out[o] = sum(f(o, s) for s in itertools.product(*map(range, sum_shape)))
    return out


def test_absorb_side_into_corner_from_left():
    """dense.py:11-14: C(0,1,2)<->T(3,4,5) -> (T0, T1, T2, C3*T6, C4*T7, C5)."""
    C, T = rnd(2, 3, 1, 2, 2, 2), rnd(3, 2, 2, 2, 3, 1, 2, 3)
    want = loop((3, 2, 2, 2, 2, 2, 3, 2), (2, 3, 1),
                lambda o, s: C[s[0], s[1], s[2], o[3], o[5], o[7]] * T[o[0], o[1], o[2], s[0], s[1], s[2], o[4], o[6]])
    want = want.reshape(3, 2, 2, 4, 6, 2)
    np.testing.assert_allclose(dense.absorb_side_into_corner_from_left(C, T), want, atol=1e-12)


def test_absorb_side_into_corner_from_right():
    """dense.py:17-20: C(3,4,5)<->T(0,1,2) -> (C0*T6, C1*T7, C2, T3, T4, T5)."""
    C, T = rnd(2, 3, 2, 2, 3, 1), rnd(2, 3, 1, 2, 2, 3, 2, 3)
    want = loop((2, 2, 3, 3, 2, 2, 2, 3), (2, 3, 1),
                lambda o, s: C[o[0], o[2], o[4], s[0], s[1], s[2]] * T[s[0], s[1], s[2], o[5], o[6], o[7], o[1], o[3]])
    want = want.reshape(4, 9, 2, 2, 2, 3)
    np.testing.assert_allclose(dense.absorb_side_into_corner_from_right(C, T), want, atol=1e-12)


@pytest.mark.parametrize("direction", range(4))
@pytest.mark.parametrize("with_op", [False, True])
def test_absorb_center_into_side(direction, with_op):
    """dense.py:23-80: T6<->S_i, T7<->S*_i, (Op0<->S*4, Op1<->S4 | S4<->S*4) ->
    (T0*S_L, T1*S*_L, T2, T3*S_R, T4*S*_R, T5, S_O, S*_O)."""
    dims = [2, 3, 1, 2]
    d = 2
    T = rnd(2, 1, 2, 1, 2, 1, dims[direction], dims[direction])
    S, Sc, Op = rnd(*dims, d), rnd(*dims, d), rnd(d, d)
    l, r, o = L(direction), R(direction), O(direction)

    def f(out, s):
        t0, sl, t1, scl, t2, t3, sr, t4, scr, t5, so, sco = out
        x, y, k, b = s
        ks, bs = [0] * 4, [0] * 4
        ks[direction], bs[direction] = x, y
        ks[l], bs[l], ks[r], bs[r], ks[o], bs[o] = sl, scl, sr, scr, so, sco
        phys = Op[b, k] * S[tuple(ks) + (k,)] * Sc[tuple(bs) + (b,)] if with_op else \
            (S[tuple(ks) + (k,)] * Sc[tuple(bs) + (k,)] if b == 0 else 0)
        return T[t0, t1, t2, t3, t4, t5, x, y] * phys
    out_shape = (2, dims[l], 1, dims[l], 2, 1, dims[r], 2, dims[r], 1, dims[o], dims[o])
    want = loop(out_shape, (dims[direction], dims[direction], d, d), f)
    want = want.reshape(2 * dims[l], dims[l], 2, dims[r], 2 * dims[r], 1, dims[o], dims[o])
    got = dense.absorb_center_sos_into_side(direction, T, S, Op, Sc) if with_op else \
        dense.absorb_center_ss_into_side(direction, T, S, Sc)
    np.testing.assert_allclose(got, want, atol=1e-12)


def test_stage1():
    """dense.py:96-99: C(3,4,5)<->T(0,1,2) -> ((C0,C1,C2), (T3,T4,T5), T6, T7)."""
    C, T = rnd(2, 1, 2, 2, 3, 1), rnd(2, 3, 1, 3, 1, 2, 2, 3)
    want = loop((2, 1, 2, 3, 1, 2, 2, 3), (2, 3, 1),
                lambda o, s: C[o[0], o[1], o[2], s[0], s[1], s[2]] * T[s[0], s[1], s[2], o[3], o[4], o[5], o[6], o[7]])
    np.testing.assert_allclose(dense.stage1(C, T), want.reshape(4, 6, 2, 3), atol=1e-12)


def test_stage2():
    """dense.py:102-112: P0<->Q1 -> (Q0, P1, P2, Q2, P3, Q3)."""
    P, Q = rnd(3, 2, 2, 3), rnd(4, 3, 3, 2)
    want = loop((4, 2, 2, 3, 3, 2), (3,), lambda o, s: P[s[0], o[1], o[2], o[4]] * Q[o[0], s[0], o[3], o[5]])
    np.testing.assert_allclose(dense.stage2(P, Q), want, atol=1e-12)


def _s2_pair():
    x, y = 2, 3
    return rnd(x, y, 2, 1, 2, 1), rnd(y, x, 1, 2, 1, 2)


@pytest.mark.parametrize("with_op", [False, True])
def test_stage3_apply(with_op):
    """dense.py:115-173 on the joined slabs: Y[a',b',c',d',s'] =
    sum_{x,y,a,b,c,d,s} s2_0[x,y,a,b,a',b'] s2_1[y,x,c,d,c',d'] Op[s',s] v[a,b,c,d,s]."""
    s0, s1 = _s2_pair()
    Op = rnd(2, 2)
    v = rnd(2, 1, 1, 2, 2)

    def f(o, s):
        x, y, a, b, c, d, k = s
        op = Op[o[4], k] if with_op else float(o[4] == k)
        return s0[x, y, a, b, o[0], o[1]] * s1[y, x, c, d, o[2], o[3]] * op * v[a, b, c, d, k]
    want = loop((2, 1, 1, 2, 2), (2, 3, 2, 1, 1, 2, 2), f)
    got = dense.stage3_apply(dense.join_a(s0), dense.join_b(s1), v, Op if with_op else None)
    np.testing.assert_allclose(got, want, atol=1e-12)


def test_stage3_matrix_and_normalization_submatrix():
    """dense.py:176-224: rows (s2_0 4,5; s2_1 4,5; Op0), cols (s2_0 2,3; s2_1 2,3; Op1); the matrix
    applied to v equals the matrix-free contractor (reference test_system.py:222-275)."""
    s0, s1 = _s2_pair()
    Op = rnd(2, 2)
    M = dense.stage3_matrix(s0, s1, Op)
    want = loop((2, 1, 1, 2, 2, 2, 1, 1, 2, 2), (2, 3),
                lambda o, s: s0[s[0], s[1], o[5], o[6], o[0], o[1]] * s1[s[1], s[0], o[7], o[8], o[2], o[3]] * Op[o[4], o[9]])
    np.testing.assert_allclose(M, want.reshape(8, 8), atol=1e-12)
    v = rnd(2, 1, 1, 2, 2)
    np.testing.assert_allclose(M @ v.ravel(), dense.stage3_apply(dense.join_a(s0), dense.join_b(s1), v, Op).ravel(), atol=1e-12)
    Nsub = dense.normalization_submatrix(s0, s1)
    np.testing.assert_allclose(np.kron(Nsub, np.identity(2)), dense.stage3_matrix(s0, s1, np.identity(2)), atol=1e-12)
    Nenv = dense.normalization_env_matrix(dense.join_a(s0), dense.join_b(s1))
    np.testing.assert_allclose(Nenv, Nsub, atol=1e-12)


def test_costs_are_the_cost_tracker_counts():
    """CostTracker (data/cost_tracker.py:17-22): prod(result shape) * prod(contracted dims) per
    pairwise step; stage 3 with an operator at D=6, chi=64, d=2 is 2 chi^2 D^6 d + D^4 d^2 MACs."""
    D, chi, d = 6, 64, 2
    c = dense.cost_stage3_multiply((chi * chi, D, D, D, D), (chi * chi, D, D, D, D), d, True)
    assert c == 2 * chi ** 2 * D ** 6 * d + D ** 4 * d ** 2
    m = dense.cost_stage3_matrix((chi, chi, D, D, D, D), (chi, chi, D, D, D, D), d)
    assert m == chi ** 2 * D ** 8 + D ** 8 * d ** 2
