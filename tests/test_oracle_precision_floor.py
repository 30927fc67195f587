"""How accurate a stage-3 matvec can be in each narrow format (CPU emulation; test infrastructure).

The low-precision kernel rounds every slab channel once (scaled to the format's range), the vector as a
hi + lo pair and the stage-A intermediate W once, and accumulates in FP32.  Emulating that rounding on the
oracle's operator (TFIM h=3, D=3, chi=12) gives the relative matvec error of each format, which bounds how
much one FP64 refinement pass of the mixed solve can gain with that format's correction operator.  The
emulated rounding is pinned against NumPy's float16 cast and the e4m3 grid's known values.  Measured
here: FP16 1.4e-5, BF16 2.9e-4, FP8 e4m3 9.8e-3 (per-channel or per-alpha scales alike), FP8 e5m2
1.4e-2..1.9e-2 on the ground state; first-order energy errors 1.4e-6, 3.4e-5, 5.6e-4, 5.5e-3..1.5e-2.
"""
import numpy as np
import pytest

FMT = {"f16": (10, -14, 65504.0), "bf16": (7, -126, 3.0e38), "e4m3": (3, -6, 448.0), "e5m2": (2, -14, 57344.0)}


def rnd(x, f):
    """Round to the nearest value of the format (ties to even), saturating at its largest finite value."""
    mant, emin, vmax = FMT[f]
    ax = np.abs(x)
    e = np.floor(np.log2(np.where(ax > 0, ax, 1.0)))
    q = 2.0 ** (np.maximum(e, emin) - mant)
    return np.clip(np.round(x / q) * q, -vmax, vmax)


def scaled_round(x, f, per_block=False):
    """Scale to the format's range (1 for the 16-bit formats as the kernel does, half the largest value for
    FP8), round, scale back; per_block: one scale per leading index (per alpha)."""
    target = 1.0 if f in ("f16", "bf16") else FMT[f][2] / 2
    if not per_block:
        s = target / np.max(np.abs(x))
        return rnd(x * s, f) / s
    m = np.max(np.abs(x.reshape(x.shape[0], -1)), axis=1).reshape((-1,) + (1,) * (x.ndim - 1))
    s = target / np.where(m > 0, m, 1.0)
    return rnd(x * s, f) / s


def emulated_apply(op, v, f, per_block=False):
    v = v.reshape(op.state_shape)
    sv = 1.0 / np.max(np.abs(v))
    hi = scaled_round(v * sv, f)
    vv = (hi + scaled_round(v * sv - hi, f)) / sv
    out = 0
    for (_, _, _, a, b, O) in op.terms:
        A, B = scaled_round(op._j(a, 0), f, per_block), scaled_round(op._j(b, 1), f, per_block)
        x = vv if O is None else np.tensordot(vv, O, (4, 1))
        W = scaled_round(np.tensordot(A, x, ((3, 4), (0, 1))), f, True)
        out = out + np.tensordot(B, W, ((0, 3, 4), (0, 3, 4))).transpose(2, 3, 0, 1, 4)
    return out.ravel()


def test_emulated_rounding_is_pinned():
    rng = np.random.default_rng(0)
    x = rng.standard_normal(10000) * 10.0 ** rng.uniform(-6, 4, 10000)
    np.testing.assert_array_equal(rnd(x, "f16"), x.astype(np.float16).astype(np.float64))
    assert rnd(np.array([0.3]), "e4m3")[0] == 0.3125           # grid spacing 2^-5 between 0.25 and 0.5
    assert rnd(np.array([1000.0]), "e4m3")[0] == 448.0          # saturates at the largest finite value
    assert rnd(np.array([2.0 ** -9]), "e4m3")[0] == 2.0 ** -9   # smallest subnormal
    assert rnd(np.array([2.0 ** -11]), "e4m3")[0] == 0.0
    assert rnd(np.array([1.0 + 2.0 ** -8]), "bf16")[0] == 1.0  # 7 stored significand bits: a tie to even


@pytest.fixture(scope="module")
def tfim_d3():
    from scipy.linalg import eigh
    from carcassonne_amd import synthetic as syn
    from oracle import sparse2d
    from tests.oracle_bridge import to_oracle
    c, s, _ = syn.make_environment(3, 12, model="ising", seed=0)
    op = sparse2d.form_multipliers(*to_oracle(c, s, syn.operator_center("ising", h=3.0)))
    H, N = op.matrix_H(), op.matrix_N()
    w, V = eigh(H, N)
    return op, H, N, w[0], V[:, 0]


def test_precision_floor_by_format(tfim_d3):
    """The matvec error grows by about 20x from FP16 to BF16 and by about 30x more to FP8 e4m3, whatever the
    FP8 scaling granularity: FP8's 1e-2 is the error of the whole correction operator, so an FP64 pass
    of the mixed solve would gain next to nothing with it (BF16, at 3e-4, already gains only ~2.5x)."""
    op, H, N, E, x = tfim_d3
    err, energy = {}, {}
    for f, blk in (("f16", False), ("bf16", False), ("e4m3", False), ("e4m3", True), ("e5m2", True)):
        y = emulated_apply(op, x, f, blk)
        err[(f, blk)] = np.linalg.norm(y - H @ x) / np.linalg.norm(H @ x)
        energy[(f, blk)] = abs((x @ y) / (x @ N @ x) - E) / abs(E)
    print({"{}{}".format(*k): "{:.2e} / {:.2e}".format(err[k], energy[k]) for k in err})
    assert err[("f16", False)] < 3e-5 and energy[("f16", False)] < 3e-6
    assert 10 * err[("f16", False)] < err[("bf16", False)] < 1e-3
    for blk in (False, True):
        assert err[("e4m3", blk)] > 10 * err[("bf16", False)] and err[("e4m3", blk)] > 3e-3
    assert err[("e5m2", True)] > err[("e4m3", True)]
