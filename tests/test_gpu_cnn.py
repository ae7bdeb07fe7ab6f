"""GPU checks of the learned start vector: the MFMA lane layout, the BF16
MFMA forward and backward kernels against the FP64 oracle, training on
engine-generated ground states, and the gain over a random start."""
import numpy as np
import pytest

import pyoracle
import spined
from spined import learned

pytestmark = pytest.mark.gpu


def test_mfma_lane_layout():
    import ctypes as C
    out = np.zeros(256, dtype=np.float32)
    spined.check_status(spined._lib.load().spined_debug_mfma_probe(out.ctypes.data))
    i = np.arange(16)[:, None]
    k = np.arange(32)[None, :]
    A = (3 * i + k) % 7 - 3
    kk = np.arange(32)[:, None]
    j = np.arange(16)[None, :]
    B = (3 * kk + 7 * j) % 5 - 2
    np.testing.assert_array_equal(out.reshape(16, 16), (A @ B).astype(np.float32))


def _bf16(x):
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


# 9: the inference kernel's smallest n; 33 and 64: windows past bit 32, all 64 bits
@pytest.mark.parametrize("n", [9, 16, 22, 33, 40, 64])
def test_forward_matches_oracle(n):
    rng = np.random.default_rng(n)
    p = learned.init_params(3)
    p[learned.NPARAMS - 34:] = rng.normal(0, 0.3, 34)   # non-trivial heads
    model = learned.CNN(params=p)
    states = np.array([sum(1 << int(i) for i in rng.choice(n, n // 2, replace=False))
                       for _ in range(100)], dtype=np.uint64)
    a, t = model.forward(n, states)
    ra, rt = pyoracle.cnn_forward(p.astype(np.float64), n, states)
    # BF16 operands with FP32 accumulation against FP64: relative to the spread
    scale = np.max(np.abs(ra)) + 1.0
    print("max |da|", np.max(np.abs(a - ra)) / scale, "max |dt|", np.max(np.abs(t - rt)) / (np.max(np.abs(rt)) + 1))
    assert np.max(np.abs(a - ra)) <= 2e-2 * scale
    assert np.max(np.abs(t - rt)) <= 2e-2 * (np.max(np.abs(rt)) + 1.0)
    assert np.corrcoef(a, ra)[0, 1] > 0.999
    # the inference path (layer 1 from a table of its 32 input patterns) and
    # the training path (layer 1 on MFMA, pre-activations stored) agree to
    # BF16 rounding
    a2, t2, _, _ = model.forward(n, states, store=True)
    assert np.max(np.abs(a - a2.numpy())) <= 5e-3 * scale
    assert np.max(np.abs(t - t2.numpy())) <= 5e-3 * (np.max(np.abs(rt)) + 1.0)


@pytest.mark.parametrize("n", [16, 26])
def test_backward_matches_oracle(n):
    rng = np.random.default_rng(100 + n)
    p = learned.init_params(5)
    p[learned.NPARAMS - 34:] = rng.normal(0, 0.3, 34)
    model = learned.CNN(params=p)
    states = np.array([sum(1 << int(i) for i in rng.choice(n, n // 2, replace=False))
                       for _ in range(70)], dtype=np.uint64)
    ga = rng.normal(size=70)
    gt = rng.normal(size=70)
    g = model.gradient(n, states, ga, gt).astype(np.float64)
    rg = pyoracle.cnn_backward(p.astype(np.float64), n, states, ga, gt)
    cos = np.dot(g, rg) / np.linalg.norm(g) / np.linalg.norm(rg)
    rel = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    print("cosine", cos, "relative error", rel)
    assert cos > 0.999
    assert rel < 5e-2
    # every parameter block is exercised
    blocks = [(0, 80), (80, 96), (96, 1376), (1376, 1392), (1392, 2672), (2672, 2688),
              (2688, 2704), (2704, 2705), (2705, 2721), (2721, 2722)]
    for b, e in blocks:
        nb = np.linalg.norm(rg[b:e])
        assert np.linalg.norm(g[b:e] - rg[b:e]) <= 0.1 * nb + 1e-6, (b, e)


def test_training_improves_start_overlap_and_lanczos():
    data = [learned.ground_state(n) for n in (12, 14, 16, 18)]
    target = learned.ground_state(22)
    model = learned.CNN(seed=11)
    ov0 = learned.overlap(model, target)
    learned.train(data, steps=300, batch=1024, lr=3e-3, model=model, log=print)
    ov = learned.overlap(model, target)
    print("overlap before", ov0, "after", ov)
    assert ov > 0.9
    k, p, z = learned.ring_sector_for_ground_state(22)
    basis, op = learned.ring_system(22, k, p, z)
    spined.build_basis(basis)
    r_rand = spined.eigsh(op, k=1, tol=1e-10)
    r_learn = spined.eigsh(op, k=1, tol=1e-10, model=model, sign_mask=learned.marshall_mask(22))
    print("matvecs random", r_rand.stats["matvecs"], "learned", r_learn.stats["matvecs"])
    assert abs(r_learn.evals[0] - r_rand.evals[0]) < 1e-10 * abs(r_rand.evals[0])
    assert r_learn.rnorms[0] <= 1e-10
    assert r_learn.stats["matvecs"] < r_rand.stats["matvecs"]
