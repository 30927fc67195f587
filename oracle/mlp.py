"""The learned warm-start network in NumPy FP64 (oracle, test infrastructure only).

No reference counterpart: the reference has no learned component; this restates the network of
carcassonne_amd/csrc/cc_mlp.hip so that its forward pass and gradients can be checked.

    x (K0) -> h1 = relu(W1 x + b1) -> h2 = relu(W2 h1 + b2) -> y = W3 h2 + b3
    loss_b = 1 - (y_b . v_b)^2 / (|y_b|^2 |v_b|^2), averaged over the batch

Parameters are one flat vector [W1 (H x K0), b1, W2 (H x H), b2, W3 (NO x H), b3].  With round16=True
the weights and the hidden activations that feed the next layer are rounded to FP16 exactly where the
GPU rounds them (the FP32 master values of the biases are used as they are).

Pinned by: the gradient against central finite differences of the loss (round16=False), and the loss
against its closed form at aligned and orthogonal outputs (tests/test_oracle_mlp.py).
"""
import numpy as np


def split_params(p, K0, H, NO):
    o = 0
    out = []
    for shape in ((H, K0), (H,), (H, H), (H,), (NO, H), (NO,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    return out


def _r16(a, on):
    return a.astype(np.float16).astype(np.float64) if on else a


def forward(p, X, K0, H, NO, round16=True):
    W1, b1, W2, b2, W3, b3 = [a.astype(np.float64) for a in split_params(p, K0, H, NO)]
    W1, W2, W3 = _r16(W1, round16), _r16(W2, round16), _r16(W3, round16)
    x = _r16(np.asarray(X, dtype=np.float64), round16)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h1r = _r16(h1, round16)
    h2 = np.maximum(h1r @ W2.T + b2, 0.0)
    h2r = _r16(h2, round16)
    y = h2r @ W3.T + b3
    return y, dict(x=x, h1=h1, h1r=h1r, h2=h2, h2r=h2r, W1=W1, W2=W2, W3=W3)


def loss(y, V):
    yy = np.sum(y * y, axis=1)
    vv = np.sum(V * V, axis=1)
    yv = np.sum(y * V, axis=1)
    return 1.0 - yv ** 2 / (yy * vv)


def backward(p, X, V, K0, H, NO, round16=True, saved=None):
    """Gradient of the mean loss with respect to the flat parameters.  `saved` may hold the hidden
    activations (h1, h1r, h2, h2r) and output y of the pass being differentiated."""
    y, s = forward(p, X, K0, H, NO, round16)
    if saved is not None:
        s.update({k: np.asarray(v, dtype=np.float64) for k, v in saved.items() if k != "y"})
        y = np.asarray(saved.get("y", y), dtype=np.float64)
    B = y.shape[0]
    yy = np.sum(y * y, axis=1, keepdims=True)
    vv = np.sum(V * V, axis=1, keepdims=True)
    yv = np.sum(y * V, axis=1, keepdims=True)
    dY = (2.0 * yv ** 2 / (yy ** 2 * vv) * y - 2.0 * yv / (yy * vv) * V) / B
    gW3 = dY.T @ s["h2r"]
    gb3 = dY.sum(axis=0)
    dZ2 = (dY @ s["W3"]) * (s["h2"] > 0)
    gW2 = dZ2.T @ s["h1r"]
    gb2 = dZ2.sum(axis=0)
    dZ1 = (dZ2 @ s["W2"]) * (s["h1"] > 0)
    gW1 = dZ1.T @ s["x"]
    gb1 = dZ1.sum(axis=0)
    return np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3)]), loss(y, V)


def features(spectra, h, K0):
    """Feature row: each spectrum divided by its largest value, then h, zero padding to K0."""
    S = np.asarray(spectra, dtype=np.float64)
    f = (S / S[:, :1]).ravel()
    out = np.zeros(K0)
    out[:f.size] = f
    out[f.size] = h
    return out
