"""The CNN oracle's FP64 backward pinned by central finite differences."""
import numpy as np

import pyoracle


def test_cnn_backward_finite_differences():
    rng = np.random.default_rng(0)
    n = 10
    p = rng.normal(0, 0.4, 2722)
    states = np.array([rng.integers(0, 1 << n) for _ in range(3)], dtype=np.uint64)
    ga, gt = rng.normal(size=3), rng.normal(size=3)
    g = pyoracle.cnn_backward(p, n, states, ga, gt)

    def L(q):
        a, t = pyoracle.cnn_forward(q, n, states)
        return np.dot(ga, a) + np.dot(gt, t)

    idx = np.concatenate([rng.choice(2722, 60, replace=False), [0, 80, 96, 1376, 2672, 2704, 2721]])
    h = 1e-5
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (L(p + e) - L(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_cnn_is_translation_invariant():
    rng = np.random.default_rng(1)
    n = 12
    p = rng.normal(0, 0.4, 2722)
    x = int(rng.integers(0, 1 << n))
    rot = ((x >> 1) | ((x & 1) << (n - 1)))
    a, t = pyoracle.cnn_forward(p, n, np.array([x, rot], dtype=np.uint64))
    assert abs(a[0] - a[1]) < 1e-12 and abs(t[0] - t[1]) < 1e-12
