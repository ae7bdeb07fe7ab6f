"""The CNN oracle pinned: its FP64 forward against an independent numpy
restatement, its backward by central finite differences of that forward,
and translation invariance."""
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


def numpy_cnn_forward(p, n, states):
    """Independent restatement of the network, vectorised over states:
    h0 = +-1 spins, three circular 5-tap convolutions of 16 channels with
    tanh-form GELU after each, sums over positions, two linear heads.  Taps
    are np.roll views (tap d reads position i + d - 2), the parameters are
    sliced from the documented layout W1[16][1][5], b1, W2[16][16][5], b2,
    W3[16][16][5], b3, u[16], u0, v[16], v0."""
    C, K = 16, 5
    def take(off, shape):
        size = int(np.prod(shape))
        return p[off:off + size].reshape(shape), off + size
    off = 0
    W1, off = take(off, (C, 1, K))
    b1, off = take(off, (C,))
    W2, off = take(off, (C, C, K))
    b2, off = take(off, (C,))
    W3, off = take(off, (C, C, K))
    b3, off = take(off, (C,))
    u, off = take(off, (C,))
    u0, off = take(off, (1,))
    v, off = take(off, (C,))
    v0, off = take(off, (1,))
    assert off == len(p)

    def gelu(x):
        return 0.5 * x * (1.0 + np.tanh(np.sqrt(2.0 / np.pi) * (x + 0.044715 * x ** 3)))

    def conv(h, W, b):  # h: (S, cin, n) -> (S, cout, n)
        taps = np.stack([np.roll(h, -(d - 2), axis=2) for d in range(K)], axis=3)  # (S, cin, n, K)
        return np.einsum("oid,sind->son", W, taps) + b[None, :, None]

    bits = (states[:, None] >> np.arange(n, dtype=np.uint64)[None, :]) & np.uint64(1)
    h = np.where(bits == 1, 1.0, -1.0)[:, None, :]
    h = gelu(conv(h, W1, b1))
    h = gelu(conv(h, W2, b2))
    f = gelu(conv(h, W3, b3)).sum(axis=2)
    return u0[0] + f @ u, v0[0] + f @ v


def test_cnn_forward_matches_numpy_restatement():
    """The oracle's forward against an independent vectorised restatement
    (random parameters with no symmetry between taps, so a wrong tap offset,
    channel order or head slice changes the result)."""
    rng = np.random.default_rng(5)
    for n in (8, 13, 24):
        p = rng.normal(0, 0.5, 2722)
        states = rng.integers(0, 1 << n, 40).astype(np.uint64)
        a, t = pyoracle.cnn_forward(p, n, states)
        ra, rt = numpy_cnn_forward(p, n, states)
        np.testing.assert_allclose(a, ra, rtol=1e-12, atol=1e-10)
        np.testing.assert_allclose(t, rt, rtol=1e-12, atol=1e-10)
