"""Dense index maps of the hot path (oracle).  Restates carcassonne/tensors/_2d/dense.py.

Leg conventions (from the join specs, not from comments):
  corner C[6]: (0,1,2) meet side L(i) legs (3,4,5); (3,4,5) meet side i legs (0,1,2)
  side   T[8]: (0,1,2) one end, (3,4,5) other end, (6,7) = center ket / bra
  state  S[5]: legs 0..3 = center bonds in directions 0..3, leg 4 = physical

Pinned by: tests/test_oracle_dense.py checks every map against an independent
explicit-loop evaluation on tiny random shapes (the reference's own strategy,
tests/test_tensors_dense.py:10-187), and the matvec against the dense matrix.
"""
import numpy as np

from .tags import L, R, O


def _merge(a, *groups):
    """Reshape consecutive axis groups (given as group sizes) into single axes."""
    shape = []
    i = 0
    for g in groups:
        shape.append(int(np.prod(a.shape[i:i + g], dtype=np.int64)))
        i += g
    assert i == a.ndim
    return a.reshape(shape)


def absorb_side_into_corner_from_left(C, T):
    """dense.py:11-14: C(0,1,2)<->T(3,4,5) -> (T0,T1,T2,(C3,T6),(C4,T7),C5)."""
    r = np.einsum("abcdef,ghiabcjk->ghidjekf", C, T)
    return _merge(r, 1, 1, 1, 2, 2, 1)


def absorb_side_into_corner_from_right(C, T):
    """dense.py:17-20: C(3,4,5)<->T(0,1,2) -> ((C0,T6),(C1,T7),C2,T3,T4,T5)."""
    r = np.einsum("abcdef,defghijk->ajbkcghi", C, T)
    return _merge(r, 2, 2, 1, 1, 1, 1)


def _state_subscripts(direction, ket, bra):
    """Subscript strings for S and S* with leg `direction` named 'x'/'y' and the
    other three legs named per their geometric role."""
    s = [None] * 4
    b = [None] * 4
    s[direction], b[direction] = "x", "y"
    s[L(direction)], b[L(direction)] = ket[0], bra[0]
    s[R(direction)], b[R(direction)] = ket[1], bra[1]
    s[O(direction)], b[O(direction)] = ket[2], bra[2]
    return "".join(s), "".join(b)


def absorb_center_ss_into_side(direction, T, S, Sc=None):
    """dense.py:23-49: T6<->S_i, T7<->S*_i, S4<->S*4 ->
    ((T0,S_L),(T1,S*_L),T2,(T3,S_R),(T4,S*_R),T5,S_O,S*_O)."""
    if Sc is None:
        Sc = S.conj()
    ks, bs = _state_subscripts(direction, "lmn", "LMN")
    r = np.einsum("abcdefxy," + ks + "z," + bs + "z->albLcdmeMfnN", T, S, Sc, optimize=True)
    return _merge(r, 2, 2, 1, 2, 2, 1, 1, 1)


def absorb_center_sos_into_side(direction, T, S, Op, Sc=None):
    """dense.py:52-80: as SS with Op0<->S*4 and Op1<->S4."""
    if Sc is None:
        Sc = S.conj()
    ks, bs = _state_subscripts(direction, "lmn", "LMN")
    r = np.einsum("abcdefxy," + ks + "z," + bs + "w,wz->albLcdmeMfnN", T, S, Sc, Op, optimize=True)
    return _merge(r, 2, 2, 1, 2, 2, 1, 1, 1)


def stage1(C, T):
    """dense.py:96-99: C(3,4,5)<->T(0,1,2) -> ((C0,C1,C2),(T3,T4,T5),T6,T7)."""
    r = np.tensordot(C, T, ((3, 4, 5), (0, 1, 2)))            # (C0,C1,C2,T3..T7)
    return _merge(r, 3, 3, 1, 1)


def stage2(P, Q):
    """dense.py:102-112: P0<->Q1 -> (Q0,P1,P2,Q2,P3,Q3)."""
    r = np.tensordot(P, Q, ((0,), (1,)))                       # (P1,P2,P3,Q0,Q2,Q3)
    return r.transpose(3, 0, 1, 4, 2, 5)


def join_a(s2_0):
    """stage2_0.join((0,1),4,5,2,3) (dense.py:130): slab layout A[alpha, bra0, bra1, ket0, ket1]."""
    t = s2_0.transpose(0, 1, 4, 5, 2, 3)
    return t.reshape((t.shape[0] * t.shape[1],) + t.shape[2:])


def join_b(s2_1):
    """stage2_1.join((1,0),4,5,2,3) (dense.py:131): slab layout B[alpha, bra2, bra3, ket2, ket3]."""
    t = s2_1.transpose(1, 0, 4, 5, 2, 3)
    return t.reshape((t.shape[0] * t.shape[1],) + t.shape[2:])


def stage3_apply(A, B, v, Op=None):
    """formNormalizationStage3 (dense.py:115-143) / formDenseStage3_multiply
    (dense.py:146-173): Y = sum_alpha A_alpha (x) B_alpha (x) Op applied to v, with A, B
    in joined slab layout.  Op=None is the normalization contractor (no physical op)."""
    if Op is not None:
        v = np.tensordot(v, Op, (4, 1))                        # (a,b,c,d,s') ; O1 <-> v4
    t = np.tensordot(A, v, ((3, 4), (0, 1)))                    # (z,p,q,c,d,s)
    r = np.tensordot(B, t, ((0, 3, 4), (0, 3, 4)))              # (r,s2,p,q,s)
    return r.transpose(2, 3, 0, 1, 4)


def stage3_matrix(s2_0, s2_1, Op):
    """formDenseStage3_formMatrix (dense.py:176-194): rows (s2_0 4,5; s2_1 4,5; Op0),
    cols (s2_0 2,3; s2_1 2,3; Op1)."""
    e = np.tensordot(s2_0, s2_1, ((0, 1), (1, 0)))             # (a,b,c,d,e,f,g,h)
    m = np.multiply.outer(e, Op)                                 # (a,b,c,d,e,f,g,h,s,t)
    m = m.transpose(2, 3, 6, 7, 8, 0, 1, 4, 5, 9)               # rows (c,d,g,h,s), cols (a,b,e,f,t)
    n = int(np.prod(m.shape[:5]))
    return m.reshape(n, n)


def normalization_submatrix(s2_0, s2_1):
    """formNormalizationSubmatrix (dense.py:205-224)."""
    e = np.tensordot(s2_0, s2_1, ((0, 1), (1, 0)))
    m = e.transpose(2, 3, 6, 7, 0, 1, 4, 5)
    n = int(np.prod(m.shape[:4]))
    return m.reshape(n, n)


# --- costs (the reference's CostTracker model, data/cost_tracker.py:17-22) -------------
def cost_stage3_multiply(A_shape, B_shape, d, with_op):
    """Complex-MAC count of the generated stage-3 contractor: first A<->v over (D0,D1),
    then B<->T over (D2,D3,alpha) (plus O<->v over d when with_op)."""
    nal, a1, a2, k0, k1 = A_shape
    _, b1, b2, k2, k3 = B_shape
    c = 0
    if with_op:
        c += (d * k0 * k1 * k2 * k3) * d
    c += (nal * a1 * a2 * k2 * k3 * d) * (k0 * k1)
    c += (b1 * b2 * a1 * a2 * d) * (k2 * k3 * nal)
    return int(c)


def cost_stage3_matrix(s2_0_shape, s2_1_shape, d):
    x, y, k0, k1, b0, b1 = s2_0_shape
    _, _, k2, k3, b2, b3 = s2_1_shape
    m = k0 * k1 * b0 * b1 * k2 * k3 * b2 * b3
    return int(m * x * y + m * d * d)


def normalization_env_matrix(AN, BN):
    """N_env[(p',q'),(p,q)] = sum_alpha A[alpha][p'][p] B[alpha][q'][q] for joined identity slabs;
    the full normalization matrix is N_env (x) I_d."""
    na, a1, a2, k0, k1 = AN.shape
    _, b1, b2, k2, k3 = BN.shape
    t = np.tensordot(AN, BN, ((0,), (0,)))                     # (a1,a2,k0,k1,b1,b2,k2,k3)
    t = t.transpose(0, 1, 4, 5, 2, 3, 6, 7)
    return t.reshape(a1 * a2 * b1 * b2, k0 * k1 * k2 * k3)
