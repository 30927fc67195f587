"""The Data protocol on HBM.  Re-implements carcassonne/data/__init__.py (NDArrayData)
for real FP64 device tensors; every arithmetic operation runs in the HIP library.

Storage is a contiguous torch tensor used only as a device buffer; shapes,
transposes and contractions are carried out by the C ABI (cc_permute,
cc_gemm_f64, cc_svd_f64, elementwise and reduction kernels).  Complex data is
rejected: the accelerated path is real (Hermitian = symmetric), which covers the
transverse-field Ising model and the Heisenberg model written with iY.
"""
import ctypes

import numpy as np
import torch

from .. import _lib
from ..device import ONES_LEN, device, empty, ones, ptr, stream

__all__ = ["Data", "HipData", "NDArrayData"]

_rng = np.random.default_rng(0)


def seed(s):
    """Seed the host generator behind newRandom / newEnlargener."""
    global _rng
    _rng = np.random.default_rng(s)


def _prod(shape):
    p = 1
    for s in shape:
        p *= int(s)
    return p


_HOST_COPY_MAX = 64   # fromArray keeps a host copy of matrices up to this many elements


def _strides(shape):
    st, acc = [], 1
    for s in reversed(shape):
        st.append(acc)
        acc *= int(s)
    return list(reversed(st))


class Data:
    def size(self):
        return _prod(self.shape)


_CONTRACT_PLANS = {}


def _contract_plan(ssh, osh, self_axes, other_axes):
    """The shape bookkeeping of one tensordot (memoised per shapes and axes): (M, N, K, output shape,
    A layout mode and permutation, B layout mode and permutation)."""
    sn, on = len(ssh), len(osh)
    self_axes = [int(a) % max(sn, 1) for a in self_axes]
    other_axes = [int(a) % max(on, 1) for a in other_axes]
    if len(self_axes) != len(other_axes):
        raise ValueError("shape-mismatch for sum")
    for a, b in zip(self_axes, other_axes):
        if ssh[a] != osh[b]:
            raise ValueError("shape-mismatch for sum")
    fa = [i for i in range(sn) if i not in self_axes]
    fb = [i for i in range(on) if i not in other_axes]
    M = _prod([ssh[i] for i in fa])
    K = _prod([ssh[i] for i in self_axes])
    N = _prod([osh[i] for i in fb])
    out_shape = [int(ssh[i]) for i in fa] + [int(osh[i]) for i in fb]
    a_mode = 0 if fa + self_axes == list(range(sn)) else (1 if self_axes + fa == list(range(sn)) else 2)
    b_mode = 0 if other_axes + fb == list(range(on)) else (1 if fb + other_axes == list(range(on)) else 2)
    return M, N, K, out_shape, a_mode, fa + self_axes, b_mode, other_axes + fb


class HipData(Data):
    """A real FP64 tensor in device memory with the NDArrayData method surface."""

    # -- construction -------------------------------------------------------------
    @classmethod
    def _wrap(cls, t):
        """A HipData around a float64 device tensor this module just produced (no input checks)."""
        d = HipData.__new__(HipData)
        d._t = t
        return d

    def __init__(self, t):
        if isinstance(t, np.ndarray) or np.isscalar(t):
            t = HipData._upload(np.asarray(t))
        if not isinstance(t, torch.Tensor):
            raise TypeError("HipData wraps a device tensor")
        if t.is_complex():
            raise TypeError("HipData is real-valued; complex tensors are not supported on the HIP path")
        if t.dtype != torch.float64 or t.device.type != "cuda":
            t = t.to(device=device(), dtype=torch.float64)
        self._t = t.contiguous()

    @staticmethod
    def _upload(a):
        if np.iscomplexobj(a):
            if np.any(a.imag != 0):
                raise TypeError("HipData is real-valued; complex tensors are not supported on the HIP path")
            a = a.real
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device())

    @classmethod
    def fromArray(cls, a):
        a = np.asarray(a)
        out = cls(cls._upload(a))
        if a.ndim == 2 and a.size <= _HOST_COPY_MAX and not np.iscomplexobj(a):
            # small operands (site operators) keep their host values: reading them back (plan building)
            # would otherwise wait for the device; in-place updates drop the copy
            out._host = np.array(a, dtype=np.float64)
        return out

    @classmethod
    def newCollected(cls, datas):
        out = empty((len(datas),) + tuple(datas[0].shape))
        for i, d in enumerate(datas):
            out[i].copy_(d._t)
        return cls(out)

    @classmethod
    def newDiagonal(cls, data):
        n = data.shape[0]
        out = cls.newZeros((n, n))
        out._t.view(-1)[:: n + 1].copy_(data._t)
        return out

    @classmethod
    def newEnlargener(cls, old_dimension, new_dimension, dtype=None):
        """data/__init__.py:43-50: a random new x old isometry (and its conjugate)."""
        if new_dimension == old_dimension:
            e = cls.newIdentity(new_dimension)
            return e, e
        m = cls.fromArray(_rng.uniform(-1, 1, (new_dimension, old_dimension)))
        U, _, Vt = m.svd(full_matrices=False)
        q = U.contractWith(Vt, (1,), (0,))
        return q, q

    @classmethod
    def newFilled(cls, shape, value, dtype=None):
        out = empty(shape)
        _lib.call("cc_fill_f64", stream(), out.numel(), float(np.real(value)), ptr(out))
        return cls(out)

    @classmethod
    def newIdentity(cls, N, dtype=None):
        out = cls.newZeros((N, N))
        _lib.call("cc_set_identity_f64", stream(), int(N), ptr(out._t))
        return out

    @classmethod
    def newOuterProduct(cls, *factors):
        r = factors[0]
        for f in factors[1:]:
            r = r.contractWith(f, (), ())
        return r

    @classmethod
    def newRandom(cls, *shape):
        """Uniform(-1, 1) entries (the real part of randomComplexSample, utils.py)."""
        return cls.fromArray(_rng.uniform(-1, 1, shape))

    @classmethod
    def newRandomHermitian(cls, *shape):
        d = cls.newRandom(*shape)
        return d + d.transpose()

    @classmethod
    def newNormalizedRandom(cls, *shape):
        d = cls.newRandom(*shape)
        return d * (1.0 / d.norm())

    @classmethod
    def newTrivial(cls, shape, dtype=None):
        return cls.newFilled(shape, 1.0)

    @classmethod
    def newZeros(cls, shape, dtype=None):
        return cls.newFilled(shape, 0.0)

    # -- properties ---------------------------------------------------------------
    shape = property(lambda self: self._t.shape)  # a torch.Size (a tuple)
    ndim = property(lambda self: self._t.dim())
    dtype = property(lambda self: np.dtype(np.float64))

    def size(self):
        return self._t.numel()

    def toArray(self):
        host = self.__dict__.get("_host")
        if host is not None:
            return host.copy()
        return self._t.cpu().numpy()

    def toNDArrayData(self):
        return self

    def tensor(self):
        return self._t

    # -- elementwise ----------------------------------------------------------------
    def _binary(self, other, op):
        if other.shape != self.shape:
            raise ValueError("operands could not be broadcast together with shapes {} {}".format(self.shape, other.shape))
        out = empty(self.shape)
        _lib.call("cc_binary_f64", stream(), op, out.numel(), ptr(self._t), ptr(other._t), ptr(out))
        return HipData(out)

    def _scaled(self, a):
        out = empty(self.shape)
        _lib.call("cc_unary_f64", stream(), 3, out.numel(), ptr(self._t), float(a), ptr(out))
        return HipData(out)

    def __add__(self, other):
        return self._binary(other, 2)

    def __sub__(self, other):
        return self._binary(other, 3)

    def __iadd__(self, other):
        self.__dict__.pop("_host", None)
        if other.shape != self.shape:
            raise ValueError("shape mismatch in +=: {} {}".format(self.shape, other.shape))
        _lib.call("cc_axpby_f64", stream(), self._t.numel(), 1.0, ptr(other._t), 1.0, ptr(self._t))
        return self

    def __imul__(self, other):
        self.__dict__.pop("_host", None)
        r = self * other
        self._t = r._t
        return self

    def __mul__(self, other):
        if isinstance(other, HipData):
            if other.shape != self.shape and other.ndim == 2 and self.ndim == 2 and other.shape[1] == 1 \
                    and other.shape[0] == self.shape[0]:
                out = empty(self.shape)
                _lib.call("cc_scale_rows_cols_f64", stream(), 0, self.shape[0], self.shape[1], ptr(self._t),
                          ptr(other._t), ptr(out))
                return HipData(out)
            return self._binary(other, 0)
        return self._scaled(float(np.real(other)))

    __rmul__ = __mul__

    def __truediv__(self, other):
        if isinstance(other, HipData):
            return self._binary(other, 1)
        return self._scaled(1.0 / float(other))

    def __neg__(self):
        return self._scaled(-1.0)

    def __copy__(self):
        return HipData(self._t.clone())

    copy = __copy__

    def __getitem__(self, index):
        return HipData(self._t[index].contiguous())

    def __setitem__(self, index, value):
        self.__dict__.pop("_host", None)
        self._t[index] = value._t

    def __repr__(self):
        return "HipData(shape={})".format(self.shape)

    def conj(self):
        return self

    def sqrt(self):
        out = empty(self.shape)
        _lib.call("cc_unary_f64", stream(), 0, out.numel(), ptr(self._t), 0.0, ptr(out))
        return HipData(out)

    # -- reductions ---------------------------------------------------------------
    def _dot(self, other):
        out = empty((1,))
        _lib.call("cc_dot_f64", stream(), self._t.numel(), ptr(self._t), ptr(other._t), ptr(out))
        return float(out.item())

    def norm(self):
        return float(np.sqrt(max(self._dot(self), 0.0)))

    def normalized(self):
        return self * (1.0 / self.norm())

    def hasNaN(self):
        out = torch.zeros(1, dtype=torch.int64, device=device())
        _lib.call("cc_count_nonfinite_f64", stream(), self._t.numel(), ptr(self._t), ptr(out))
        return bool(out.item() > 0)

    def allcloseTo(self, other, rtol=1e-05, atol=1e-08):
        a, b = self.toArray(), other.toArray()
        return bool(np.allclose(a, b, rtol=rtol, atol=atol))

    def isCloseTo(self, other, rtol=1e-7, atol=1e-7):
        nd = (self - other).norm()
        return nd <= atol or nd / (self.norm() + other.norm()) / 2 <= rtol

    def extractScalar(self):
        if self.ndim != 0:
            raise ValueError("tensor is not a scalar")
        return float(self._t.item())

    # -- shape operations (cc_permute) ----------------------------------------------------
    def _permuted(self, perm):
        perm = [int(p) for p in perm]
        sh = self.shape
        if perm == list(range(len(sh))):
            return self._t
        shape = [sh[p] for p in perm]
        st = _strides(sh)
        out = empty(shape)
        if out.numel():
            _lib.call("cc_permute", stream(), _lib.CC_F64, ptr(self._t), ptr(out), len(perm),
                      _lib.i64_array(shape), _lib.i64_array([st[p] for p in perm]))
        return out

    def strided_source(self):
        """(contiguous device tensor, element strides of this tensor's legs in it)."""
        return self._t, _strides(self.shape)

    def permuted_lazily(self, perm):
        """This tensor with its legs permuted, materialised only when a kernel needs the permuted data."""
        perm = [int(p) for p in perm]
        if perm == list(range(self.ndim)):
            return self
        return PermutedView(self._t, perm)

    def transpose(self, *args):
        if len(args) == 1 and not isinstance(args[0], (int, np.integer)):
            args = tuple(args[0])
        perm = list(args) if args else list(reversed(range(self.ndim)))
        return HipData(self._permuted(perm))

    def join(self, *groups):
        groups = [[g] if isinstance(g, (int, np.integer)) else list(g) for g in groups]
        perm = [i for g in groups for i in g]
        t = self._permuted(perm)
        shape = [_prod([self.shape[i] for i in g]) for g in groups]
        return HipData(t.reshape(shape))

    def split(self, *splits):
        return HipData(self._t.reshape(tuple(int(s) for s in splits)))

    def splitAt(self, index, *split):
        shape = list(self.shape)
        assert _prod(split) == shape[index]
        return self.split(*(shape[:index] + list(split) + shape[index + 1:]))

    def ravel(self):
        return HipData(self._t.reshape(-1))

    def fold(self, axis):
        others = [i for i in range(self.ndim) if i != axis]
        return self.join(axis, others)

    def dropUnitAxis(self, axis):
        if self.shape[axis] != 1:
            raise ValueError("Axis {} has non-unit dimension {}.".format(axis, self.shape[axis]))
        return self.split(*[s for i, s in enumerate(self.shape) if i != axis])

    def reverseLastAxis(self):
        return HipData(torch.flip(self._t, dims=(-1,)))

    def directSumWith(self, other, *non_summed_axes):
        if self.ndim != other.ndim:
            raise ValueError("In a direct sum the number of axes must match")
        new_shape, li, ri = [], [], []
        for axis, (m, n) in enumerate(zip(self.shape, other.shape)):
            if axis in non_summed_axes:
                if m != n:
                    raise ValueError("All non-summed axes must agree in shape")
                new_shape.append(m); li.append(slice(0, m)); ri.append(slice(0, m))
            else:
                new_shape.append(m + n); li.append(slice(0, m)); ri.append(slice(m, m + n))
        out = HipData.newZeros(new_shape)
        out._t[tuple(li)] = self._t
        out._t[tuple(ri)] = other._t
        return out

    # -- contraction (cc_gemm_f64) ----------------------------------------------------
    def contractWith(self, other, self_axes, other_axes, split_long_sums=False):
        """tensordot(self, other, (self_axes, other_axes)) on the GPU.  split_long_sums: a product with a
        small output and a long sum is summed in slices (_gemm); the sum's rounding then differs from the
        single product's, so only callers that ask for it get it."""
        ssh, osh = self.shape, other.shape
        self_axes, other_axes = tuple(self_axes), tuple(other_axes)
        key = (ssh, osh, self_axes, other_axes)
        plan = _CONTRACT_PLANS.get(key)
        if plan is None:
            plan = _contract_plan(ssh, osh, self_axes, other_axes)
            if len(_CONTRACT_PLANS) < 4096:
                _CONTRACT_PLANS[key] = plan
        M, N, K, out_shape, a_mode, a_perm, b_mode, b_perm = plan
        out = empty(out_shape)
        if out.numel() == 0:
            return HipData._wrap(out)
        if K == 0:
            _lib.call("cc_fill_f64", stream(), out.numel(), 0.0, ptr(out))
            return HipData._wrap(out)
        # A as (M, K): free-then-contracted (0) or contracted-then-free (1) in memory order, else permuted
        if a_mode == 0:
            A, sAm, sAk = self._t, K, 1
        elif a_mode == 1:
            A, sAm, sAk = self._t, 1, M
        else:
            A, sAm, sAk = self._permuted(a_perm), K, 1
        if b_mode == 0:
            B, sBk, sBn = other._t, N, 1
        elif b_mode == 1:
            B, sBk, sBn = other._t, 1, K
        else:
            B, sBk, sBn = other._permuted(b_perm), N, 1
        if split_long_sums:
            _gemm(M, N, K, A, sAm, sAk, B, sBk, sBn, out)
        else:
            _lib.call("cc_gemm_f64", stream(), 1, M, N, K, 1.0, ptr(A), 0, sAm, sAk, ptr(B), 0, sBk, sBn, 0.0,
                      ptr(out), 0, N, 1)
        return HipData._wrap(out)

    def contractWithAlongAll(self, other):
        assert self.ndim == other.ndim
        return self.contractWith(other, range(self.ndim), range(self.ndim))

    def absorbMatrixAt(self, axis, matrix):
        """data/__init__.py:148-150: new[..., i, ...] = sum_j matrix[i, j] self[..., j, ...]."""
        return matrix.contractWith(self, (1,), (axis,)).join(*(tuple(range(1, axis + 1)) + (0,) + tuple(range(axis + 1, self.ndim))))

    def matvecWith(self, v):
        return v.absorbMatrixAt(0, self)

    def adjoint(self):
        if self.ndim != 2:
            raise ValueError("Adjoint may only be computed for rank 2 tensors.")
        return self.join(1, 0)

    # -- factorizations (cc_svd_f64) ----------------------------------------------------
    def svd(self, full_matrices=True):
        """Thin SVD on the GPU (one-sided Jacobi).  full_matrices is honoured only when
        it does not change the result (square or thin requests)."""
        if self.ndim != 2:
            raise ValueError("svd needs a matrix")
        m, n = self.shape
        if m < n:
            Ut, S, Vtt = HipData(self._t.t().contiguous()).svd(False)
            return Vtt.join(1, 0), S, Ut.join(1, 0)
        U, S, Vt = empty((m, n)), empty((n,)), empty((n, n))
        ws = _lib.i64_array([0])
        _lib.call("cc_svd_work_size", 1, m, n, ws)
        work = empty((max(1, ws[0]),))
        _lib.call("cc_svd_f64", stream(), 1, m, n, ptr(self._t), ptr(U), ptr(S), ptr(Vt), ptr(work))
        return HipData(U), HipData(S), HipData(Vt)

    def qr(self, mode="full"):
        """Q with orthonormal columns spanning the column space (via the SVD), and R = Q^T A."""
        U, S, Vt = self.svd(False)
        Q = U
        return Q, Q.contractWith(self, (0,), (0,))

    def unitize(self):
        """utils.py:879-882: the polar factor U Vt."""
        U, _, Vt = self.svd(False)
        return U.contractWith(Vt, (1,), (0,))

    def normalizeAxis(self, axis, sqrt_svals=False, dont_recip_under=1e-14):
        """data/__init__.py:263-301."""
        if self.shape[axis] == 1:
            n = self.norm()
            if sqrt_svals:
                n = np.sqrt(n)
                return HipData.fromArray(np.array([[1 / n]])), HipData.fromArray(np.array([[n]]))
            return self * (1.0 / n), HipData.fromArray(np.array([[1 / n]])), HipData.fromArray(np.array([[n]]))
        others = [i for i in range(self.ndim) if i != axis]
        M = self.join(others, axis)
        if M.shape[0] < M.shape[1]:
            raise ValueError("the total number of degrees of freedom in all other axes ({}) are not enough to "
                             "normalize axis ({}) with dimension ({})".format(M.shape[0], axis, M.shape[1]))
        U, S, V = M.svd(full_matrices=False)
        U = U.contractWith(V, (1,), (0,))
        S = S.split(S.shape[0], 1)
        SI = HipData(empty(S.shape))
        if dont_recip_under:
            _lib.call("cc_unary_f64", stream(), 1, S.size(), ptr(S._t), float(dont_recip_under), ptr(SI._t))
        else:
            _lib.call("cc_unary_f64", stream(), 4, S.size(), ptr(S._t), 0.0, ptr(SI._t))
        if sqrt_svals:
            S, SI = S.sqrt(), SI.sqrt()
            return (V * SI), V * S
        U_split = [self.shape[i] for i in others] + [U.shape[1]]
        U_join = list(range(self.ndim - 1))
        U_join.insert(axis, self.ndim - 1)
        return (U.split(*U_split).join(*U_join), V.transpose().contractWith(V * SI, (1,), (0,)),
                V.transpose().contractWith(V * S, (1,), (0,)))

    def normalizeAxisAndDenormalize(self, axis_to_norm, axis_to_denorm, data_to_denormalize=None, sqrt_svals=False,
                                    dont_recip_under=1e-14):
        if data_to_denormalize is None:
            data_to_denormalize = self
        if self.shape[axis_to_norm] != data_to_denormalize.shape[axis_to_denorm]:
            raise ValueError("Normalized axis and denormalized axis have different sizes ({} != {}).".format(
                self.shape[axis_to_norm], data_to_denormalize.shape[axis_to_denorm]))
        normalized, _, denormalizer = self.normalizeAxis(axis_to_norm, sqrt_svals, dont_recip_under)
        return normalized, data_to_denormalize.absorbMatrixAt(axis_to_denorm, denormalizer)


NDArrayData = HipData


class _DeviceConstant:
    """A class-level constant matrix uploaded on first use (NDArrayData.I/X/Z,
    data/__init__.py:360-363).  Y is complex; the real path uses iY = [[0,1],[-1,0]]."""

    def __init__(self, rows, name):
        self.rows, self.name, self.value = rows, name, None

    def __get__(self, obj, cls):
        if self.rows is None:
            raise TypeError("NDArrayData.{} is complex; use HipData.iY (= i*Y) on the real HIP path".format(self.name))
        if self.value is None:
            self.value = HipData.fromArray(np.array(self.rows, dtype=np.float64))
        return self.value


HipData.I = _DeviceConstant([[1, 0], [0, 1]], "I")
HipData.X = _DeviceConstant([[0, 1], [1, 0]], "X")
HipData.Z = _DeviceConstant([[1, 0], [0, -1]], "Z")
HipData.iY = _DeviceConstant([[0, 1], [-1, 0]], "iY")
HipData.Y = _DeviceConstant(None, "Y")


def _gemm(M, N, K, A, sAm, sAk, B, sBk, sBn, out):
    """out (M x N, row-major) = A (M x K, strides sAm, sAk) B (K x N, strides sBk, sBn).  A product with a
    small output and a long sum (M and N from 32 to 64, K >= 2048: the corner projections of
    the sweep step at D=6, chi=64 sum over 2304) would occupy a handful of workgroups for its whole
    length; its sum is cut into up to 64 slices of at least 32 (one batched launch, plus one for the
    remainder) and the slices are added in order by a second product with a vector of ones."""
    if not (32 <= M <= 64 and 32 <= N <= 64 and K >= 2048):
        _lib.call("cc_gemm_f64", stream(), 1, M, N, K, 1.0, ptr(A), 0, sAm, sAk, ptr(B), 0, sBk, sBn, 0.0,
                  ptr(out), 0, N, 1)
        return
    S = min(64, K // 64)
    kc = K // S
    rest = K - S * kc
    parts = empty((S + (1 if rest else 0), M, N))
    _lib.call("cc_gemm_f64", stream(), S, M, N, kc, 1.0, ptr(A), kc * sAk, sAm, sAk, ptr(B), kc * sBk, sBk, sBn,
              0.0, ptr(parts), M * N, N, 1)
    if rest:
        a_rest = ptr(A).value + S * kc * sAk * 8
        b_rest = ptr(B).value + S * kc * sBk * 8
        _lib.call("cc_gemm_f64", stream(), 1, M, N, rest, 1.0, ctypes.c_void_p(a_rest), 0, sAm, sAk,
                  ctypes.c_void_p(b_rest), 0, sBk, sBn, 0.0, ptr(parts[S]), 0, N, 1)
    n_parts = parts.shape[0]
    assert n_parts <= ONES_LEN
    _lib.call("cc_gemm_f64", stream(), 1, 1, M * N, n_parts, 1.0, ptr(ones()), 0, n_parts, 1, ptr(parts), 0, M * N, 1,
              0.0, ptr(out), 0, M * N, 1)


class PermutedView(HipData):
    """A leg permutation of a contiguous device tensor that is carried out (cc_permute) only when some
    kernel needs the permuted data as a contiguous array.  Readers that take strides (the stage-3 slab
    packers) use strided_source() and skip the copy.  Constructed with a single tensor it is a plain
    HipData (so classmethods called on type(x) keep working)."""

    def __init__(self, base, perm=None):
        if perm is None:
            HipData.__init__(self, base)
            return
        if not (isinstance(base, torch.Tensor) and base.is_contiguous() and base.dtype == torch.float64):
            raise TypeError("PermutedView needs a contiguous float64 device tensor")
        self._base, self._perm, self._mat = base, [int(p) for p in perm], None
        bs = base.shape
        self._shape = torch.Size([bs[p] for p in self._perm])

    @property
    def _t(self):
        if self._mat is None:
            self._mat = HipData(self._base)._permuted(self._perm)
            self._base = None
        return self._mat

    @_t.setter
    def _t(self, value):
        self._mat, self._base = value, None
        self._shape = value.shape

    def _lazy(self):
        return self.__dict__.get("_mat") is None and self.__dict__.get("_base") is not None

    @property
    def shape(self):
        return self._shape

    @property
    def ndim(self):
        return len(self.shape)

    def size(self):
        return _prod(self.shape)

    def strided_source(self):
        if self._lazy():
            st = _strides(tuple(self._base.shape))
            return self._base, [st[p] for p in self._perm]
        return self._t, _strides(self.shape)

    def _same_lazy_layout(self, other):
        return (self._lazy() and isinstance(other, PermutedView) and other._lazy() and other._perm == self._perm and
                tuple(other._base.shape) == tuple(self._base.shape))

    def __iadd__(self, other):
        # sums of contraction results with the same leg permutation stay unpermuted (sparse accumulation)
        if self._same_lazy_layout(other):
            _lib.call("cc_axpby_f64", stream(), self._base.numel(), 1.0, ptr(other._base), 1.0, ptr(self._base))
            return self
        return HipData.__iadd__(self, other)

    def __add__(self, other):
        if self._same_lazy_layout(other):
            out = empty(tuple(self._base.shape))
            _lib.call("cc_binary_f64", stream(), 2, out.numel(), ptr(self._base), ptr(other._base), ptr(out))
            return PermutedView(out, self._perm)
        return HipData.__add__(self, other)

