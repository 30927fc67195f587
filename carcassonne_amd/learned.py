"""The learned warm start: a small network that proposes the eigensolver's start vector.

No reference counterpart: the reference always starts relaxOver from the current center
(carcassonne/system/_2d.py:491-492), which is the "previous-step" start.  Here a 3-layer MLP
(hidden 512, FP16 weights, FP32 accumulation) maps pooled environment features -- the singular
spectra of the four corners and four sides (identity channel), each divided by its largest value,
plus the field h -- to the D^4 d center tensor.  It is trained on (environment, converged state)
pairs that the plain FP64 solve produces, with the HIP forward / backward / Adam kernels of
cc_mlp.hip.  The plain solve stays the arbiter: a proposal is only ever a start vector, the solve
still runs to its own tolerance.

Every numerical step runs in the HIP library; torch tensors are buffers only.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .data import HipData
from .device import empty, ptr, stream, synchronize


def _round16(x):
    return (x + 15) // 16 * 16


def spectra_of_environment(corners, sides):
    """Device (8, chi) FP64 singular values: four corners as (chi x chi) matrices and four sides as
    ((chi D D) x chi) matricizations of their identity channel."""
    from .sparse import Identity
    cs = [c[Identity()]._t for c in corners]
    ss = [s[Identity()] for s in sides]
    chi = cs[0].shape[0]
    C = torch.stack([c.reshape(chi, chi) for c in cs]).contiguous()
    # rows (y, a, b), columns x: the side's outward bond against its inward bond and center legs
    S = torch.stack([s.join((3, 4, 5, 6, 7), (0, 1, 2))._t for s in ss]).contiguous()
    m_side = S.shape[1]
    out = empty((8, chi))
    for batch, (A, m) in enumerate(((C, chi), (S, m_side))):
        U, Vt = empty((4, m, chi)), empty((4, chi, chi))
        work = empty((_lib.query("cc_svd_work_size", 4, m, chi),))
        _lib.call("cc_svd_f64", stream(), 4, m, chi, ptr(A), ptr(U), ptr(out[4 * batch:4 * batch + 4]), ptr(Vt),
                  ptr(work))
    return out


class WarmStartMLP:
    """x (K0) -> relu(W1 x + b1) (H) -> relu(W2 . + b2) (H) -> W3 . + b3 (NO)."""

    def __init__(self, n_spec, k, n_out, hidden=512, seed=0):
        self.n_spec, self.k = n_spec, k
        self.K0, self.H, self.NO = _round16(n_spec * k + 1), hidden, n_out
        if hidden % 64 or n_out % 64:
            raise ValueError("hidden width and output size must be multiples of 64")
        self.n_params = _lib.query("cc_mlp_param_count", self.K0, self.H, self.NO)
        rng = np.random.Generator(np.random.PCG64(seed))
        p = np.zeros(self.n_params, dtype=np.float32)
        off = 0
        for n_in, n_out_l in ((self.K0, self.H), (self.H, self.H), (self.H, self.NO)):
            lim = np.sqrt(6.0 / n_in)                      # He-uniform for the ReLU layers
            p[off:off + n_out_l * n_in] = rng.uniform(-lim, lim, n_out_l * n_in)
            off += n_out_l * n_in + n_out_l               # bias stays zero
        self.P32 = torch.from_numpy(p).to(empty((1,)).device)
        self.P16 = torch.empty(self.n_params, dtype=torch.float16, device=self.P32.device)
        self._refresh16()
        self.M = torch.zeros_like(self.P32)
        self.V = torch.zeros_like(self.P32)
        self.G = torch.zeros_like(self.P32)
        self.step = 0

    def _refresh16(self):
        _lib.call("cc_cast", stream(), self.n_params, _lib.CC_F32, ptr(self.P32), _lib.CC_F16, ptr(self.P16))

    @classmethod
    def for_residual(cls, n, n_out, hidden=512, seed=0):
        """A network whose input is a residual row (residual_features: n values and h) instead of spectra."""
        return cls(1, n, n_out, hidden=hidden, seed=seed)

    # -- features ---------------------------------------------------------------------------
    def features(self, corners, sides, h, out=None):
        """FP16 feature row (K0) of one environment."""
        S = spectra_of_environment(corners, sides)
        if out is None:
            out = torch.empty(self.K0, dtype=torch.float16, device=S.device)
        _lib.call("cc_mlp_features", stream(), 8, S.shape[1], ptr(S), float(h), self.K0, ptr(out))
        return out

    # -- forward / backward ---------------------------------------------------------------------
    def _buffers(self, B):
        dev = self.P32.device
        A16 = torch.empty(2 * B * self.H, dtype=torch.float16, device=dev)
        A32 = torch.empty(2 * B * self.H, dtype=torch.float32, device=dev)
        Y = torch.empty((B, self.NO), dtype=torch.float32, device=dev)
        return A16, A32, Y

    def forward(self, X16):
        """X16: (B, K0) FP16 device tensor -> (B, NO) FP32 outputs (plus the saved activations)."""
        B = X16.shape[0]
        A16, A32, Y = self._buffers(B)
        _lib.call("cc_mlp_forward", stream(), B, self.K0, self.H, self.NO, ptr(self.P16), ptr(self.P32), ptr(X16),
                  ptr(A16), ptr(A32), ptr(Y))
        return Y, (A16, A32)

    def loss_and_grad(self, X16, target):
        """Mean misalignment loss of a batch and its gradient (into self.G)."""
        B = X16.shape[0]
        Y, (A16, A32) = self.forward(X16)
        loss = torch.empty(B, dtype=torch.float32, device=Y.device)
        work = torch.empty(_lib.query("cc_mlp_work_size", B, self.H, self.NO), dtype=torch.float32, device=Y.device)
        _lib.call("cc_mlp_backward", stream(), B, self.K0, self.H, self.NO, ptr(self.P16), ptr(X16), ptr(A16),
                  ptr(A32), ptr(Y), ptr(target), ptr(loss), ptr(self.G), ptr(work))
        return loss, Y

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        _lib.call("cc_mlp_adam", stream(), self.n_params, ptr(self.P32), ptr(self.G), ptr(self.M), ptr(self.V),
                  ptr(self.P16), float(lr), float(beta1), float(beta2), float(eps), self.step)

    def train(self, X16, targets, epochs=20, batch=256, lr=1e-3, seed=0, lr_decay=0.9):
        """Minibatch Adam over (X16, targets) (device tensors, rows = samples); returns the mean loss per
        epoch.  The sample order is drawn from a seeded host generator."""
        n = X16.shape[0]
        rng = np.random.Generator(np.random.PCG64(seed))
        history = []
        for ep in range(epochs):
            order = torch.from_numpy(rng.permutation(n)).to(X16.device)
            losses = []
            for b0 in range(0, n - batch + 1, batch):
                idx = order[b0:b0 + batch]
                xb, tb = X16.index_select(0, idx).contiguous(), targets.index_select(0, idx).contiguous()
                loss, _ = self.loss_and_grad(xb, tb)
                self.adam_step(lr * lr_decay ** ep)
                losses.append(loss)
            history.append(float(torch.cat(losses).mean().item()) if losses else float("nan"))
        return history

    def propose(self, x16_row):
        """Start vector (FP64 device, NO) for one feature row."""
        Y, _ = self.forward(x16_row.reshape(1, -1).contiguous())
        out = empty((self.NO,))
        _lib.call("cc_cast", stream(), self.NO, _lib.CC_F32, ptr(Y), _lib.CC_F64, ptr(out))
        return out

    # -- persistence ------------------------------------------------------------------------------
    def state(self):
        return dict(P32=self.P32.cpu().numpy(), dims=np.array([self.n_spec, self.k, self.NO, self.H]))

    @classmethod
    def from_state(cls, state):
        n_spec, k, NO, H = (int(v) for v in state["dims"])
        m = cls(n_spec, k, NO, hidden=H)
        m.P32.copy_(torch.from_numpy(np.asarray(state["P32"], dtype=np.float32)))
        m._refresh16()
        return m


# -- data and evaluation on environment families ---------------------------------------------------
def _pad64(n):
    return (n + 63) // 64 * 64


class Family:
    """Environments of one sweep-like family: a base environment (base_seed) plus a seeded
    perturbation of relative size sigma (carcassonne_amd.synthetic.make_environment), with field h."""

    def __init__(self, D, chi, model="ising", base_seed=0, sigma=0.2):
        self.D, self.chi, self.model, self.base_seed, self.sigma = D, chi, model, base_seed, sigma
        self.n = D ** 4 * 2

    def environment(self, sample_seed, h):
        from . import synthetic as syn
        from .pipeline import device_environment
        c, s, st = syn.make_environment(self.D, self.chi, model=self.model, seed=self.base_seed,
                                        perturb_seed=sample_seed, sigma=self.sigma)
        corners, sides, opc = device_environment(c, s, syn.operator_center(self.model, h=h))
        return corners, sides, opc, HipData.fromArray(st)


def combined_start(H_operator, N_operator, vectors):
    """The Rayleigh-Ritz combination of several start vectors (flat FP64 device tensors): the lowest
    generalized eigenvector of the pencil (V^T H V, V^T N V) over their span, as one flat vector.  With the
    previous step's state and the network's proposal it is never worse, in energy, than either alone; its
    cost is one FP64 H and N product per vector (cc_small_geneig_f64 on the device)."""
    k = len(vectors)
    n = H_operator.n
    V = torch.stack([v.reshape(-1) for v in vectors]).contiguous()
    HV, NV = torch.empty_like(V), torch.empty_like(V)
    for j in range(k):
        H_operator.apply_flat(V[j], HV[j])
        N_operator.apply_flat(V[j], NV[j])
    hs, ns = empty((k, k)), empty((k, k))
    for j in range(k):
        _lib.call("cc_basis_dots_f64", stream(), k, n, ptr(V), ptr(HV[j]), ptr(hs[:, j]), k)
        _lib.call("cc_basis_dots_f64", stream(), k, n, ptr(V), ptr(NV[j]), ptr(ns[:, j]), k)
    small = empty((k + 1,))
    _lib.call("cc_small_geneig_f64", stream(), k, ptr(hs), ptr(ns), ptr(small[0:1]), ptr(small[1:1 + k]))
    coeffs = small[1:1 + k].contiguous()
    out = empty((n,))
    _lib.call("cc_basis_combine_f64", stream(), k, n, ptr(V), ptr(coeffs), ptr(out))
    return out


def residual_features(net, H_operator, N_operator, x_prev, h, out=None):
    """FP16 input row of the correction network: the residual r = H x - rho N x of the previous step's state
    x in the new environment (rho its Rayleigh quotient), scaled to unit norm, then h, zero-padded to K0.  The
    residual carries what the new environment changes about the state (to first order the correction is
    -(H - rho N)^-1 r off x), which the environment's spectra alone do not."""
    n = H_operator.n
    if net.K0 < n + 1:
        raise ValueError("the network's input is narrower than the residual")
    if out is None:
        out = torch.zeros(net.K0, dtype=torch.float16, device=net.P32.device)
    hx, nx, r = empty((n,)), empty((n,)), empty((n,))
    H_operator.apply_flat(x_prev, hx)
    N_operator.apply_flat(x_prev, nx)
    d = empty((3,))
    _lib.call("cc_dot_f64", stream(), n, ptr(x_prev), ptr(hx), ptr(d[0:1]))
    _lib.call("cc_dot_f64", stream(), n, ptr(x_prev), ptr(nx), ptr(d[1:2]))
    _lib.call("cc_binary_f64", stream(), 1, 1, ptr(d[0:1]), ptr(d[1:2]), ptr(d[2:3]))      # rho
    _lib.call("cc_axpby_dev_f64", stream(), n, ptr(d[2:3]), -1.0, ptr(nx), 1.0, ptr(hx))    # hx -= rho nx
    _lib.call("cc_normalize_f64", stream(), n, ptr(hx), ptr(r))
    _lib.call("cc_cast", stream(), n, _lib.CC_F64, ptr(r), _lib.CC_F16, ptr(out[0:n]))
    hv = torch.full((1,), float(h), dtype=torch.float64, device=out.device)
    _lib.call("cc_cast", stream(), 1, _lib.CC_F64, ptr(hv), _lib.CC_F16, ptr(out[n:n + 1]))
    return out


def learned_correction_start(net, H_operator, N_operator, x_prev, h):
    """The learned start of a sweep step: the network maps the previous state's residual in the new
    environment to a correction direction y, and the start is the Rayleigh-Ritz combination of the previous
    state and y (combined_start)."""
    y = net.propose(residual_features(net, H_operator, N_operator, x_prev, h))[:H_operator.n].contiguous()
    return combined_start(H_operator, N_operator, [x_prev, y])


def generate_correction_pairs(family, net, n_chains, rng, h_start=(2.0, 2.3), h_step=(0.04, 0.13), seed0=1000,
                              check_every=4):
    """Training pairs of the correction network from walks that mimic consecutive sweep steps: each chain
    starts at h ~ U[h_start] with a fresh perturbation seed and steps h by U[h_step] (new perturbation seed
    every step) until h > 4.  For each step: the input row from the previous state's residual in the new
    environment, and the target, the part of the new converged state orthogonal to the previous one (sign
    aligned with it).  Returns (X16, T, number of pairs)."""
    from .pipeline import multipliers
    n = family.n
    dev = net.P32.device
    X, T = [], []
    seed = seed0
    for c in range(n_chains):
        h = float(rng.uniform(*h_start))
        corners, sides, opc, x0 = family.environment(seed, h)
        seed += 1
        x_prev, _, _ = solve_from(corners, sides, opc, x0, check_every=check_every)
        while True:
            h += float(rng.uniform(*h_step))
            if h > 4.0:
                break
            corners, sides, opc, _ = family.environment(seed, h)
            seed += 1
            H, N = multipliers(corners, sides, opc)
            X.append(residual_features(net, H.operator, N.operator, x_prev, h))
            x, _, _ = solve_from(corners, sides, opc, x_prev, check_every=check_every)
            d = empty((2,))
            _lib.call("cc_dot_f64", stream(), n, ptr(x), ptr(x_prev), ptr(d[0:1]))
            _lib.call("cc_dot_f64", stream(), n, ptr(x_prev), ptr(x_prev), ptr(d[1:2]))
            _lib.call("cc_binary_f64", stream(), 1, 1, ptr(d[0:1]), ptr(d[1:2]), ptr(d[0:1]))
            t = x.clone()
            _lib.call("cc_axpby_dev_f64", stream(), n, ptr(d[0:1]), -1.0, ptr(x_prev), 1.0, ptr(t))   # t = x - c x_prev
            tt = torch.zeros(net.NO, dtype=torch.float32, device=dev)
            _lib.call("cc_cast", stream(), n, _lib.CC_F64, ptr(t), _lib.CC_F32, ptr(tt))
            T.append(tt)
            x_prev = x
    return torch.stack(X).contiguous(), torch.stack(T).contiguous(), len(X)


def solve_from(corners, sides, op_center, x0, check_every=1, tol_residual=1e-10):
    """Plain FP64 solve from start vector x0 (flat device tensor or HipData): (x, energy, iterations)."""
    from .pipeline import solve
    r = solve(corners, sides, op_center, x0, tol_residual=tol_residual, tol_energy=0.0, check_every=check_every)
    return r["x"], r["energy"], r["iterations"]


def generate_pairs(family, model_net, sample_seeds, h_values, check_every=4):
    """Features (n, K0) FP16 and converged states (n, NO) FP32 (zero-padded to the network's output size)
    from plain solves started at the family's seeded random state."""
    n = len(sample_seeds)
    dev = model_net.P32.device
    X16 = torch.empty((n, model_net.K0), dtype=torch.float16, device=dev)
    T = torch.zeros((n, model_net.NO), dtype=torch.float32, device=dev)
    iters = []
    for i, (sd, h) in enumerate(zip(sample_seeds, h_values)):
        corners, sides, opc, x0 = family.environment(sd, h)
        x, _, it = solve_from(corners, sides, opc, x0, check_every=check_every)
        model_net.features(corners, sides, h, out=X16[i])
        _lib.call("cc_cast", stream(), family.n, _lib.CC_F64, ptr(x), _lib.CC_F32, ptr(T[i]))
        iters.append(it)
    return X16, T, iters


def evaluate_starts(family, model_net, sample_seeds, h_values, seed=0, correction_net=None):
    """Lanczos iterations to a 1e-10 residual on each held-out environment from four starts: a seeded
    random vector, the previous environment's converged state (the sequence is walked in order, as
    consecutive sweep steps), the network's proposal, and the Rayleigh-Ritz combination of the last two
    (combined_start; timed with its products).  Energies of the solves agree."""
    import time
    rng = np.random.Generator(np.random.PCG64(seed))
    out = dict(random=[], previous=[], learned=[], combined=[], corrected=[], energy_spread=[], ms_random=[],
               ms_previous=[], ms_learned=[], ms_combined=[], ms_corrected=[], overlap_learned=[],
               overlap_previous=[], overlap_combined=[], overlap_corrected=[])

    def overlap(a, b):
        """|a . b|^2 / (|a|^2 |b|^2) on the device (cc_dot_f64)."""
        d = empty((3,))
        _lib.call("cc_dot_f64", stream(), family.n, ptr(a), ptr(b), ptr(d[0:1]))
        _lib.call("cc_dot_f64", stream(), family.n, ptr(a), ptr(a), ptr(d[1:2]))
        _lib.call("cc_dot_f64", stream(), family.n, ptr(b), ptr(b), ptr(d[2:3]))
        ab, aa, bb = d.tolist()
        return ab * ab / (aa * bb)
    prev = None

    def timed(fn):
        synchronize()
        t0 = time.perf_counter()
        r = fn()
        synchronize()
        return r, (time.perf_counter() - t0) * 1e3

    for sd, h in zip(sample_seeds, h_values):
        corners, sides, opc, _ = family.environment(sd, h)
        xr = torch.from_numpy(rng.standard_normal(family.n)).to(model_net.P32.device)
        (x1, e1, it_r), ms = timed(lambda: solve_from(corners, sides, opc, xr))
        out["random"].append(it_r)
        out["ms_random"].append(ms)

        def learned():
            xl = model_net.propose(model_net.features(corners, sides, h))[:family.n].contiguous()
            return solve_from(corners, sides, opc, xl)
        (_, e2, it_l), ms = timed(learned)
        out["learned"].append(it_l)
        out["ms_learned"].append(ms)
        xl = model_net.propose(model_net.features(corners, sides, h))[:family.n].contiguous()
        out["overlap_learned"].append(overlap(xl, x1))
        es = [e1, e2]
        if prev is not None:
            (_, e3, it_p), ms = timed(lambda: solve_from(corners, sides, opc, prev))
            out["previous"].append(it_p)
            out["ms_previous"].append(ms)
            out["overlap_previous"].append(overlap(prev, x1))
            es.append(e3)

            def combined():
                from .pipeline import multipliers
                H, N = multipliers(corners, sides, opc)
                xl = model_net.propose(model_net.features(corners, sides, h))[:family.n].contiguous()
                xc = combined_start(H.operator, N.operator, [prev, xl])
                return solve_from(corners, sides, opc, xc) + (xc,)
            (_, e4, it_c, xc), ms = timed(combined)
            out["combined"].append(it_c)
            out["ms_combined"].append(ms)
            out["overlap_combined"].append(overlap(xc, x1))
            es.append(e4)
            if correction_net is not None:
                def corrected():
                    from .pipeline import multipliers
                    H, N = multipliers(corners, sides, opc)
                    xs = learned_correction_start(correction_net, H.operator, N.operator, prev, h)
                    return solve_from(corners, sides, opc, xs) + (xs,)
                (_, e5, it_x, xs), ms = timed(corrected)
                out["corrected"].append(it_x)
                out["ms_corrected"].append(ms)
                out["overlap_corrected"].append(overlap(xs, x1))
                es.append(e5)
        out["energy_spread"].append(max(es) - min(es))
        prev = x1
    return out
