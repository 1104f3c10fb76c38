"""Pins the FP64 oracle of the proposal network (CPU only).

* the dilated convolution equals a brute-force loop and, at dilation 1,
  oracle.conv_f64 (itself pinned against main/fprop.cu's loop);
* the analytic backward pass equals central finite differences of the loss
  for every layer's weights and biases;
* the Huber loss has its closed form at known points.
"""
import numpy as np

import oracle
from oracle import proposer_ref as P

LAYERS = [(2, 4, 1, True), (4, 4, 2, True), (4, 2, 3, False)]


def test_dilated_conv_matches_loops():
    g = np.random.Generator(np.random.PCG64(1))
    x = g.standard_normal((2, 3, 7, 9))
    w = g.standard_normal((4, 3, 3, 3))
    b = g.standard_normal(4)
    for d in (1, 2, 3):
        out = P.conv(x, w, b, d)
        ref = np.zeros_like(out)
        for n in range(2):
            for o in range(4):
                for yy in range(7):
                    for xx in range(9):
                        s = b[o]
                        for c in range(3):
                            for r in range(3):
                                for q in range(3):
                                    y2, x2 = yy + (r - 1) * d, xx + (q - 1) * d
                                    if 0 <= y2 < 7 and 0 <= x2 < 9:
                                        s += w[o, c, r, q] * x[n, c, y2, x2]
                        ref[n, o, yy, xx] = s
        np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(P.conv(x, w, np.zeros(4), 1), oracle.conv_f64(x, w, 1, 1), rtol=1e-12, atol=1e-12)


def test_backward_matches_finite_differences():
    g = np.random.Generator(np.random.PCG64(2))
    x = g.standard_normal((2, 2, 8, 12))
    params = [(g.standard_normal((co, ci, 3, 3)) * 0.5, g.standard_normal(co) * 0.1) for ci, co, _, _ in LAYERS]
    labels = g.integers(-8, 9, size=(2, 4, 6, 3))

    def loss_of(ps):
        acts = P.forward(x, ps, LAYERS)
        return P.huber(acts[-1], labels)[0]

    acts = P.forward(x, params, LAYERS)
    _, gout = P.huber(acts[-1], labels)
    grads = P.backward(x, params, LAYERS, acts, gout)
    eps = 1e-6
    for li in range(len(LAYERS)):
        for which in (0, 1):
            arr = params[li][which]
            for idx in [tuple(g.integers(0, s) for s in arr.shape) for _ in range(6)]:
                plus = [(w.copy(), b.copy()) for w, b in params]
                minus = [(w.copy(), b.copy()) for w, b in params]
                plus[li][which][idx] += eps
                minus[li][which][idx] -= eps
                fd = (loss_of(plus) - loss_of(minus)) / (2 * eps)
                assert abs(fd - grads[li][which][idx]) <= 1e-6 + 1e-5 * abs(fd), (li, which, idx)


def test_huber_closed_form():
    pred = np.zeros((1, 2, 2, 2))
    labels = np.zeros((1, 1, 1, 3), np.int64)
    labels[0, 0, 0, :2] = (2, 12)  # targets 0.5 and 3.0
    loss, grad = P.huber(pred, labels)
    assert np.isclose(loss, (4 * 0.125 + 4 * 2.5) / 8)
    assert np.allclose(grad[0, 0], -0.5 / 8) and np.allclose(grad[0, 1], -1.0 / 8)
