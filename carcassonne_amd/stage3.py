"""The center-site effective-Hamiltonian operator of one solve.

Re-implements formNormalizationStage3 / formDenseStage3 / multiplyExpectation
(carcassonne/tensors/_2d/dense.py:115-203, carcassonne/tensors/_2d/sparse.py:129-148):

    (H v)[p', q', s'] = sum_t sum_alpha sum_{p,q,s} A_t[alpha][p'][p] B_t[alpha][q'][q] O_t[s'][s] v[p, q, s]

with A_t = stage2_0.join((0,1),4,5,2,3) and B_t = stage2_1.join((1,0),4,5,2,3) -- the
alpha-major slab layout -- and O_t the center operator (identity for the
normalization contractor).  All terms of a solve are applied in one call to the
HIP library: the fused MFMA kernel for the instantiated slab shapes, the generic
GEMM path for any other shape.
"""
import ctypes
import weakref

import numpy as np

from . import _lib
from .data import HipData
from .device import empty, ptr, stream


def cost_of_term(n_alpha, P1, P, Q1, Q, d, with_op):
    """The reference's CostTracker count (complex MACs) of one stage-3 contractor:
    [O<->v], then A<->v over (D0,D1), then B<->T over (D2,D3,alpha)."""
    c = (d * P * Q) * d if with_op else 0
    c += (n_alpha * P1 * Q * d) * P
    c += (Q1 * P1 * d) * (Q * n_alpha)
    return int(c)


def cost_of_matrix(s2_0_shape, s2_1_shape, d):
    x, y, k0, k1, b0, b1 = s2_0_shape
    _, _, k2, k3, b2, b3 = s2_1_shape
    m = k0 * k1 * b0 * b1 * k2 * k3 * b2 * b3
    return int(m * x * y + m * d * d)


class _Slab:
    """One slab channel of a stage-3 operator: the stage-2 tensor t read as slab[alpha][p'][p] with
    which 0: alpha = (t0, t1) (stage2_0.join((0,1),4,5,2,3)), which 1: alpha = (t1, t0)
    (stage2_1.join((1,0),4,5,2,3)); p' = (t4, t5), p = (t2, t3).  The FP64 packing for the fused kernel
    and the low-precision conversion read t in place through strides; the joined (alpha, P, P) array
    is formed only on request (generic path, normalization matrix).  One _Slab per (tensor, which),
    cached on the tensor, so operators built from the same stage-2 tensors share packed channels."""

    def __init__(self, t, which):
        # the channel is cached on t: hold t only weakly (no reference cycle, so the stage-2 tensor and
        # its packed copies are freed as soon as the last operator using them goes)
        self._tref, self.which = weakref.ref(t), which
        self._base, self._st = t.strided_source()
        self.shape = tuple(int(x) for x in t.shape)
        if len(self.shape) != 6:
            raise ValueError("a stage-2 tensor has six legs, not {}".format(len(self.shape)))
        a0, a1 = (0, 1) if which == 0 else (1, 0)
        self.na0, self.na1 = self.shape[a0], self.shape[a1]
        self.n_alpha = self.na0 * self.na1
        self.P = self.shape[2] * self.shape[3]
        self.strided = self.shape[2] == self.shape[4] and self.shape[3] == self.shape[5]
        self._joined, self._packed, self.lowprec, self._packed_f32 = None, None, {}, {}

    @staticmethod
    def of(t, which):
        cache = t.__dict__.setdefault("_stage3_slabs", {})
        if which not in cache:
            cache[which] = _Slab(t, which)
        return cache[which]

    @property
    def t(self):
        t = self._tref()
        if t is None:
            raise RuntimeError("the stage-2 tensor of this slab channel no longer exists")
        return t

    def source(self):
        """(base tensor, (na0, na1, d0, d1), six strides) of the slab map."""
        st = self._st
        a0, a1 = (0, 1) if self.which == 0 else (1, 0)
        return self._base, (self.na0, self.na1, self.shape[4], self.shape[5]), [st[a0], st[a1], st[4], st[5], st[2], st[3]]

    @property
    def joined(self):
        if self._joined is None:
            if self.strided:
                base, (na0, na1, d0, d1), st = self.source()
                out = empty((self.n_alpha, d0 * d1, self.P))
                _lib.call("cc_permute", stream(), _lib.CC_F64, ptr(base), ptr(out), 6,
                          _lib.i64_array([na0, na1, d0, d1, d0, d1]), _lib.i64_array(st))
                self._joined = HipData(out)
            else:
                t = self.t
                self._joined = t.join((0, 1), 4, 5, 2, 3) if self.which == 0 else t.join((1, 0), 4, 5, 2, 3)
        return self._joined

    @property
    def _t(self):
        return self.joined._t

    def packed(self):
        """The channel in the fused kernel's MFMA-operand order (cc_stage3_pack_strided_f64)."""
        if self._packed is None:
            buf = empty((_lib.query("cc_stage3_packed_size", self.n_alpha, self.P),))
            base, (na0, na1, d0, d1), st = self.source()
            _lib.call("cc_stage3_pack_strided_f64", stream(), na0, na1, d0, d1, _lib.i64_array(st), ptr(base), ptr(buf))
            self._packed = buf
        return self._packed


    def packed_f32(self, role):
        """The channel in the FP32 kernel's operand order (cc_stage3_pack_strided_f32): role 0 as an A
        channel, role 1 as a B channel."""
        if role not in self._packed_f32:
            import torch
            buf = torch.empty(_lib.query("cc_stage3_packed_size_f32", self.n_alpha, self.P, role), dtype=torch.float32,
                              device=self._base.device)
            base, (na0, na1, d0, d1), st = self.source()
            _lib.call("cc_stage3_pack_strided_f32", stream(), role, na0, na1, d0, d1, _lib.i64_array(st), ptr(base),
                      ptr(buf))
            self._packed_f32[role] = buf
        return self._packed_f32[role]


class Stage3Operator:
    """terms: list of (stage2_0, stage2_1, site_operator or None) with HipData tensors."""

    def __init__(self, terms, physical_dimension=None):
        self.terms = list(terms)
        if not self.terms:
            raise ValueError("a stage-3 operator needs at least one term")
        s0, s1, o = self.terms[0]
        self.d = int(physical_dimension if physical_dimension is not None else
                     next((op.shape[0] for (_, _, op) in self.terms if op is not None), None))
        self.state_shape = (s0.shape[2], s0.shape[3], s1.shape[2], s1.shape[3], self.d)
        self.n = int(np.prod(self.state_shape))
        self.P, self.Q = s0.shape[2] * s0.shape[3], s1.shape[2] * s1.shape[3]
        self.n_alpha = s0.shape[0] * s0.shape[1]
        self._A, self._B, self._O = [], [], []
        for (a, b, op) in self.terms:
            if a.shape[0] * a.shape[1] != self.n_alpha or b.shape[0] * b.shape[1] != self.n_alpha:
                raise ValueError("stage-2 tensors of one operator must share the environment bond")
            self._A.append(self._join(a, 0))
            self._B.append(self._join(b, 1))
            self._O.append(op)
        self.cost_of_multiply = sum(cost_of_term(self.n_alpha, a.shape[4] * a.shape[5], self.P, b.shape[4] * b.shape[5],
                                                 self.Q, self.d, op is not None)
                                    for a, b, op in zip(self._A, self._B, self._O))
        self.cost_of_formMatrix = sum(cost_of_matrix(a.shape, b.shape, self.d) for (a, b, _) in self.terms)
        self._work = None
        self._fused = None
        self.use_fused = (self.P == self.Q and all(x.strided for x in self._A + self._B) and
                          bool(_lib.load().cc_stage3_fused_supported(self.P, self.d)) and
                          self._plan_fits())

    def _plan_fits(self):
        na = len({id(a) for a in self._A})
        nb = len({id(b) for b in self._B})
        nops = len({id(o) for o in self._O if o is not None})
        return na <= 16 and nb <= 16 and len(self._A) <= 48 and nops <= 16

    def _channels(self):
        """Distinct A and B channels in first-use order, the distinct site operators (HipData), and the
        term table (A index, term, B index, operator index) sorted by A channel."""
        a_ids, b_ids, o_ids = {}, {}, {}
        a_list, b_list, ops = [], [], []
        rows = []
        for k, (A, B, O) in enumerate(zip(self._A, self._B, self._O)):
            if id(A) not in a_ids:
                a_ids[id(A)] = len(a_list)
                a_list.append(A)
            if id(B) not in b_ids:
                b_ids[id(B)] = len(b_list)
                b_list.append(B)
            o = -1
            if O is not None:
                if id(O) not in o_ids:
                    o_ids[id(O)] = len(ops)
                    ops.append(O)
                o = o_ids[id(O)]
            rows.append((a_ids[id(A)], k, b_ids[id(B)], o))
        rows.sort()
        return a_list, b_list, ops, rows

    def refresh_fused_from(self, other):
        """Packs the channels of `other` -- an operator of the same structure (the same terms, channel
        sharing and shapes) built from new stage-2 tensors -- into this operator's packed buffers, so that
        its fused kernel and native plan apply `other` (kernels only; capturable in a HIP graph)."""
        self._ensure_fused()
        a_list, b_list, _, rows = self._channels()
        oa, ob, _, orows = other._channels()
        if ([(r[0], r[2], r[3]) for r in rows] != [(r[0], r[2], r[3]) for r in orows] or
                [(c.n_alpha, c.P) for c in a_list + b_list] != [(c.n_alpha, c.P) for c in oa + ob]):
            raise ValueError("refresh_fused_from needs an operator of the same structure")
        for ch, buf in zip(oa + ob, self._fused["A"] + self._fused["B"]):
            base, (na0, na1, d0, d1), st = ch.source()
            _lib.call("cc_stage3_pack_strided_f64", stream(), na0, na1, d0, d1, _lib.i64_array(st), ptr(base), ptr(buf))

    def _build_fused(self):
        """Pack every distinct slab channel once and build the term table (sorted by A channel)."""
        a_list, b_list, ops, rows = self._channels()
        packs = [ch.packed() for ch in a_list + b_list]
        self._fused = dict(
            A=packs[:len(a_list)], B=packs[len(a_list):],
            A_ptrs=_lib.ptr_array([t.data_ptr() for t in packs[:len(a_list)]]),
            B_ptrs=_lib.ptr_array([t.data_ptr() for t in packs[len(a_list):]]),
            ta=_lib.i32_array([r[0] for r in rows]), tb=_lib.i32_array([r[2] for r in rows]),
            to=_lib.i32_array([r[3] for r in rows]), n_terms=len(rows), n_a=len(a_list), n_b=len(b_list),
            n_ops=len(ops), ops=_lib.f64_array([x for o in ops for x in o.toArray().reshape(-1)]),
            work=empty((max(1, _lib.query("cc_stage3_fused_work_size", self.n_alpha, self.P, self.d)),)))

    def _join(self, t, which):
        return _Slab.of(t, which)

    def _ensure_fused(self):
        if self._fused is None:
            self._build_fused()

    def plan_handle(self, precision="f64"):
        """A native plan (cc_stage3_plan_create) for solvers that apply this operator many times; precision
        "f32" gives the FP32-operand plan (cc_stage3_plan_create_f32) of the same operator."""
        self._ensure_fused()
        f = self._fused
        if precision == "f32":
            if f.get("plan_f32") is None:
                a_list, b_list, _, _ = self._channels()
                pa = [ch.packed_f32(0) for ch in a_list]
                pb = [ch.packed_f32(1) for ch in b_list]
                f["f32_slabs"] = pa + pb
                h = ctypes.c_void_p()
                _lib.call("cc_stage3_plan_create_f32", ctypes.byref(h), f["n_a"], _lib.ptr_array([t.data_ptr() for t in pa]),
                          f["n_b"], _lib.ptr_array([t.data_ptr() for t in pb]), f["n_terms"], f["ta"], f["tb"], f["to"],
                          f["n_ops"], f["ops"], self.n_alpha, self.P, self.d, ptr(f["work"]))
                f["plan_f32"] = h
            return f["plan_f32"]
        if precision != "f64":
            raise ValueError("precision must be 'f64' or 'f32', not {!r}".format(precision))
        if f.get("plan") is None:
            h = ctypes.c_void_p()
            _lib.call("cc_stage3_plan_create", ctypes.byref(h), f["n_a"], f["A_ptrs"], f["n_b"], f["B_ptrs"],
                      f["n_terms"], f["ta"], f["tb"], f["to"], f["n_ops"], f["ops"], self.n_alpha, self.P, self.d,
                      ptr(f["work"]))
            f["plan"] = h
        return f["plan"]

    def __del__(self):
        f = getattr(self, "_fused", None)
        for key in ("plan", "plan_f32"):
            if f and f.get(key) is not None:
                try:
                    _lib.load().cc_stage3_plan_destroy(f[key])
                except Exception:
                    pass

    def apply_flat(self, x, y):
        """y = H x for flat device tensors of length n (torch float64)."""
        if self.use_fused:
            self._ensure_fused()
            f = self._fused
            _lib.call("cc_stage3_matvec_fused_f64", stream(), f["n_a"], f["A_ptrs"], f["n_b"], f["B_ptrs"],
                      f["n_terms"], f["ta"], f["tb"], f["to"], f["n_ops"], f["ops"], self.n_alpha, self.P, self.d,
                      ptr(x), ptr(y), ptr(f["work"]))
            return
        self.apply_generic_flat(x, y)

    def apply_generic_flat(self, x, y):
        """The reference's contraction order through the generic GEMM path (any shape)."""
        if self._work is None:
            ws = _lib.i64_array([0])
            _lib.call("cc_stage3_matvec_generic_work_size", self.n_alpha, self.P, self.Q, self.d, ws)
            self._work = empty((max(1, ws[0]),))
        _lib.call("cc_stage3_matvec_generic_f64", stream(), len(self._A),
                  _lib.ptr_array([a._t.data_ptr() for a in self._A]),
                  _lib.ptr_array([b._t.data_ptr() for b in self._B]),
                  _lib.ptr_array([o._t.data_ptr() if o is not None else 0 for o in self._O]),
                  self.n_alpha, self.P, self.Q, self.d, ptr(x), ptr(y), ptr(self._work))

    def apply(self, v):
        if tuple(v.shape) != self.state_shape and v.size() != self.n:
            raise ValueError("state shape {} does not match the operator's {}".format(v.shape, self.state_shape))
        out = empty(self.state_shape)
        self.apply_flat(v._t.reshape(-1), out.view(-1))
        return HipData(out)

    def form_matrix(self):
        from .tensors._2d.dense import _form_matrix
        eye = HipData.newIdentity(self.d)
        m = None
        for (a, b, op) in self.terms:
            t = _form_matrix(a, b, op if op is not None else eye)
            m = t if m is None else m + t
        return m

    def multiplier(self):
        from .utils import Multiplier
        m = Multiplier((self.n, self.n), self.apply, self.cost_of_multiply, self.form_matrix, self.cost_of_formMatrix)
        m.operator = self
        return m
