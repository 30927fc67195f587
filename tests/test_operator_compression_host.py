"""Host check of the product's operator-channel compression logic (System.compressCornerTwoSiteOperatorTowards,
carcassonne/system/_2d.py:229-363) on a NumPy-backed Data stand-in against the oracle's restatement: the
same channel dictionaries in, the same compressed channels out (up to the compressor's row gauge, so the
check is on gauge-invariant corner-side products)."""
import numpy as np

from carcassonne_amd import sparse as PT
from oracle import tags as OT


class NumpyData:
    def __init__(self, a):
        self.a = np.asarray(a, dtype=float)

    shape = property(lambda self: self.a.shape)
    ndim = property(lambda self: self.a.ndim)

    @classmethod
    def newCollected(cls, datas):
        return cls(np.array([d.a for d in datas]))

    @classmethod
    def newZeros(cls, shape):
        return cls(np.zeros(shape))

    @classmethod
    def fromArray(cls, a):
        return cls(a)

    def toArray(self):
        return self.a

    def conj(self):
        return self

    def dropUnitAxis(self, axis):
        assert self.a.shape[axis] == 1
        return NumpyData(np.squeeze(self.a, axis=axis))

    def join(self, *groups):
        groups = [[g] if isinstance(g, (int, np.integer)) else list(g) for g in groups]
        t = self.a.transpose([i for g in groups for i in g])
        shape, k = [], 0
        for g in groups:
            shape.append(int(np.prod(t.shape[k:k + len(g)], dtype=np.int64)))
            k += len(g)
        return NumpyData(t.reshape(shape))

    def fold(self, axis):
        return self.join(axis, [i for i in range(self.ndim) if i != axis])

    def contractWith(self, other, sa, oa):
        return NumpyData(np.tensordot(self.a, other.a, (list(sa), list(oa))))

    def directSumWith(self, other):
        m = np.zeros((self.shape[0] + other.shape[0], self.shape[1] + other.shape[1]))
        m[:self.shape[0], :self.shape[1]] = self.a
        m[self.shape[0]:, self.shape[1]:] = other.a
        return NumpyData(m)

    def svd(self, full_matrices=True):
        u, s, vt = np.linalg.svd(self.a, full_matrices=False)
        return NumpyData(u), NumpyData(s), NumpyData(vt)

    def __getitem__(self, idx):
        return NumpyData(self.a[idx])

    def transpose(self, *args):
        if len(args) == 1 and not isinstance(args[0], (int, np.integer)):
            args = tuple(args[0])
        return NumpyData(self.a.transpose(args))

    def absorbMatrixAt(self, axis, m):
        return NumpyData(np.moveaxis(np.tensordot(m.a, self.a, (1, axis)), 0, axis))

    def __add__(self, o):
        return NumpyData(self.a + o.a)

    def __mul__(self, o):
        return NumpyData(self.a * (o.a if isinstance(o, NumpyData) else o))

    def __truediv__(self, o):
        return NumpyData(self.a / o.a)


class _Holder:
    pass


def test_product_compression_logic_matches_oracle():
    from carcassonne_amd.system._2d import System as PSystem
    from oracle.system import System as OSystem
    rng = np.random.default_rng(2)
    corner_id, direction = 0, 0
    n1, new = 4, 3
    axis, saxis = 3 * direction + 2, 3 * (1 - direction) + 2
    cshape = [2, 3, 2, 3, 2, 2]
    cshape[axis] = 1
    sshape = [3, 2, 2, 3, 2, 2, 2, 2]
    sshape[saxis] = 1
    corner_o = {OT.Identity(): rng.standard_normal(cshape)}
    side_o = {OT.Identity(): rng.standard_normal(sshape)}
    for p in range(n1):
        corner_o[OT.TwoSiteOperator(0, direction, p)] = rng.standard_normal(cshape)
        side_o[OT.TwoSiteOperator(0, 1 - direction, p)] = rng.standard_normal(sshape)
    side_id = (corner_id + 1 - direction) % 4
    osys = OSystem.__new__(OSystem)
    osys.corners = [dict(corner_o) if i == corner_id else {} for i in range(4)]
    osys.sides = [dict(side_o) if i == side_id else {} for i in range(4)]
    osys.compressCornerTwoSiteOperatorTowards(corner_id, direction, new)
    conv = lambda t: PT.Identity() if isinstance(t, OT.Identity) else PT.TwoSiteOperator(t.id, t.direction, t.position)
    psys = _Holder()
    psys.corners = [{conv(t): NumpyData(v) for t, v in corner_o.items()} if i == corner_id else {} for i in range(4)]
    psys.sides = [{conv(t): NumpyData(v) for t, v in side_o.items()} if i == side_id else {} for i in range(4)]
    PSystem.compressCornerTwoSiteOperatorTowards(psys, corner_id, direction, new)
    oc = osys.corners[corner_id][OT.TwoSiteOperatorCompressed(direction)]
    os_ = osys.sides[side_id][OT.TwoSiteOperatorCompressed(1 - direction)]
    pc = psys.corners[corner_id][PT.TwoSiteOperatorCompressed(direction)].a
    ps = psys.sides[side_id][PT.TwoSiteOperatorCompressed(1 - direction)].a
    assert pc.shape == oc.shape and ps.shape == os_.shape
    # gauge-invariant: sum over the compressed channel of corner (x) side
    pair = lambda c, s: np.tensordot(np.moveaxis(c, axis, 0), np.moveaxis(s, saxis, 0), (0, 0))
    np.testing.assert_allclose(pair(pc, ps), pair(oc, os_), rtol=1e-10, atol=1e-12)


def test_compute_compressor_rank_deficient_gram_uses_orthogonal_rows():
    """A rank-1 Gram: the compressor rows stay orthonormal (they come from the rotation factor)."""
    from carcassonne_amd.utils import Multiplier, computeCompressor
    v = np.array([1.0, 2.0, -1.0, 0.5])
    M = NumpyData(np.outer(v, v))
    C, Ci = computeCompressor(4, 3, Multiplier((4, 4), None, 16, lambda: M, 0))
    np.testing.assert_allclose(C.a @ C.a.T, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(abs(C.a[0] @ v) / np.linalg.norm(v), 1.0, atol=1e-12)
