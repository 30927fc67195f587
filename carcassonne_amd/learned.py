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


def evaluate_starts(family, model_net, sample_seeds, h_values, seed=0):
    """Lanczos iterations to a 1e-10 residual on each held-out environment from three starts: a seeded
    random vector, the previous environment's converged state (the sequence is walked in order, as
    consecutive sweep steps), and the network's proposal.  Energies of the three solves agree."""
    import time
    rng = np.random.Generator(np.random.PCG64(seed))
    out = dict(random=[], previous=[], learned=[], energy_spread=[], ms_random=[], ms_previous=[], ms_learned=[],
               overlap_learned=[], overlap_previous=[])

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
        out["energy_spread"].append(max(es) - min(es))
        prev = x1
    return out
