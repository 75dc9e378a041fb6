"""numpy statement of the learned shader's network, for checking the HIP kernels.

The network has no counterpart in the reference; these functions state what the
kernels are specified to compute: a 12 -> 64 -> 64 -> 3 MLP (ReLU, ReLU, linear) whose
inputs, weights and hidden activations are rounded to bfloat16 (round to nearest
even) and whose products are accumulated in at least FP32.  ``grad`` is the gradient
of the mean over a batch of the squared error summed over channels, with the
rounding treated as the identity (straight-through), which is how the training
kernel differentiates.
"""
import numpy as np

N_IN, N_HID, N_OUT = 12, 64, 3
N_PARAMS = N_HID * N_IN + N_HID + N_HID * N_HID + N_HID + N_OUT * N_HID + N_OUT   # 5187


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.asarray(x, np.float32)
    u = np.ascontiguousarray(a.reshape(-1)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def _rounder(rounded):
    """bf16 rounding of float32 values, or (unrounded) float64 as given."""
    if rounded:
        return lambda a: bf16(np.asarray(a, np.float32)).astype(np.float64)
    return lambda a: np.asarray(a, np.float64)


def unpack(params):
    p = np.asarray(params)
    o = 0
    out = []
    for shape in ((N_HID, N_IN), (N_HID,), (N_HID, N_HID), (N_HID,), (N_OUT, N_HID), (N_OUT,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    assert o == N_PARAMS
    return out


def forward(params, x, rounded=True):
    """(n, 12) -> (n, 3) in float64 arithmetic on the bf16-rounded operands."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    r = _rounder(rounded)
    X = r(x)
    h1 = r(np.maximum(X @ r(W1).T + b1, 0.0))
    h2 = r(np.maximum(h1 @ r(W2).T + b2, 0.0))
    return h2 @ r(W3).T + b3


def loss_and_grad(params, x, target, rounded=True):
    """Mean over rows of the summed squared error, and its gradient (straight-through)."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    r = _rounder(rounded)
    X = r(x)
    w1, w2, w3 = (r(w) for w in (W1, W2, W3))
    a1 = X @ w1.T + b1
    h1 = r(np.maximum(a1, 0.0))
    a2 = h1 @ w2.T + b2
    h2 = r(np.maximum(a2, 0.0))
    y = h2 @ w3.T + b3
    n = X.shape[0]
    res = y - np.asarray(target, np.float64)
    loss = float((res * res).sum() / n)
    dy = 2.0 * res / n
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    d2 = (dy @ w3) * (a2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ w2) * (a1 > 0)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def random_params(seed: int, scale: float = 1.0) -> np.ndarray:
    """He-uniform weights and small random biases."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (N_HID, N_IN)) * np.sqrt(6 / N_IN)
    W2 = rng.uniform(-1, 1, (N_HID, N_HID)) * np.sqrt(6 / N_HID)
    W3 = rng.uniform(-1, 1, (N_OUT, N_HID)) * np.sqrt(6 / N_HID)
    b1, b2, b3 = (rng.uniform(-0.1, 0.1, n) for n in (N_HID, N_HID, N_OUT))
    p = np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3]) * scale
    return p.astype(np.float32)
