"""CPU pins of the numpy statement of the learned shader (tests/mlp_ref.py): bf16
rounding against a brute-force nearest-even search, the gradient against central
finite differences, and the forward pass against an explicit per-row loop."""
import numpy as np

import mlp_ref as M


def _bf16_brute(x: np.float32) -> np.float32:
    """Nearest bfloat16 to x by exact comparison of the two bracketing candidates; ties
    go to the even mantissa."""
    u = int(np.float32(x).view(np.uint32))
    lo = u & 0xFFFF0000
    hi = lo + 0x10000
    a = float(np.uint32(lo).view(np.float32))
    b = float(np.uint32(hi).view(np.float32))
    xv = float(x)
    da, db = abs(xv - a), abs(b - xv)
    if da < db:
        return np.uint32(lo).view(np.float32)
    if db < da:
        return np.uint32(hi).view(np.float32)
    return np.uint32(lo if ((lo >> 16) & 1) == 0 else hi).view(np.float32)


def test_bf16_known_answers():
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5, 0.0, -0.0, 3.0e-39],
                    np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, 0.0, -0.0, None], object)
    got = M.bf16(vals)
    for g, w in zip(got[:-1], want[:-1]):
        assert g == np.float32(w)
    assert np.signbit(got[6])


def test_bf16_matches_brute_force():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(3000).astype(np.float32) * 10,
                        (rng.random(1000).astype(np.float32) - 0.5) * 1e-30,
                        # exact ties: low half 0x8000 with odd and even upper mantissas
                        (np.arange(64, dtype=np.uint32) * 0x10000 + 0x3F808000).view(np.float32)])
    got = M.bf16(x)
    for a, g in zip(x, got):
        assert g.view(np.uint32) == _bf16_brute(a).view(np.uint32), a


def test_forward_matches_row_loop():
    p = M.random_params(3)
    x = np.random.default_rng(4).uniform(-1, 1, (17, M.N_IN)).astype(np.float32)
    y = M.forward(p, x)
    W1, b1, W2, b2, W3, b3 = (np.asarray(a, np.float64) for a in M.unpack(p))
    for i in range(x.shape[0]):
        xi = [float(M.bf16(np.float32(v))) for v in x[i]]
        h1 = []
        for o in range(M.N_HID):
            s = sum(float(M.bf16(np.float32(W1[o, k]))) * xi[k] for k in range(M.N_IN)) + b1[o]
            h1.append(float(M.bf16(np.float32(max(s, 0.0)))))
        h2 = []
        for o in range(M.N_HID):
            s = sum(float(M.bf16(np.float32(W2[o, k]))) * h1[k] for k in range(M.N_HID)) + b2[o]
            h2.append(float(M.bf16(np.float32(max(s, 0.0)))))
        for c in range(M.N_OUT):
            s = sum(float(M.bf16(np.float32(W3[c, k]))) * h2[k] for k in range(M.N_HID)) + b3[c]
            assert abs(s - y[i, c]) < 1e-9


def test_gradient_matches_finite_differences():
    p = M.random_params(5).astype(np.float64)
    rng = np.random.default_rng(6)
    x = rng.uniform(-1, 1, (9, M.N_IN))
    t = rng.uniform(0, 1, (9, M.N_OUT))
    _, g = M.loss_and_grad(p, x, t, rounded=False)
    idx = np.concatenate([rng.choice(M.N_PARAMS, 300, replace=False),
                          np.arange(M.N_PARAMS - 3, M.N_PARAMS)])
    h = 1e-6
    for e in idx:
        pp, pm = p.copy(), p.copy()
        pp[e] += h
        pm[e] -= h
        fd = (M.loss_and_grad(pp, x, t, rounded=False)[0] - M.loss_and_grad(pm, x, t, rounded=False)[0]) / (2 * h)
        assert abs(fd - g[e]) <= 1e-5 * (1 + abs(g[e])), (e, fd, g[e])
