"""Block matvec (row-interleaved K vectors, spined_operator_matmat_rows_device
and the block path of ls_operator_matmat) against the single-vector kernel
and the CPU oracle.

Bars: per column the block kernel does the same FP64 sums in the same order
as the single-vector kernel, so columns are bit-identical to it; both agree
with the oracle within 1e-13 of max |y|.
"""
import numpy as np
import pytest

import spined
from spined import lattice
from spined.device import DeviceArray
from helpers import Spec, engine, oracle, ring

pytestmark = pytest.mark.gpu

SIGMA_XX = [[0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0], [1, 0, 0, 0]]

BLOCK_CASES = [
    ring(20, k=0, p=0, z=1),                                   # exchange bonds, ring orbits
    ring(18, k=9, p=1, z=-1),                                  # sign characters
    Spec(16, 8, 1, lattice.square_symmetries(4),
         [(lattice.HEISENBERG, lattice.square_bonds(4))], "square4x4"),
    ring(12, k=0, p=0, z=None, w=None, matrix=SIGMA_XX, name="xx12_templates"),
]

F32, F64 = 0, 1  # LS_FLOAT32, LS_FLOAT64


def _rows_matvec(op, D, X, dtype_code, y_is_fp64=False):
    K = X.shape[1]
    xd = DeviceArray.from_numpy(np.ascontiguousarray(X))       # row-interleaved
    ydt = np.float64 if (dtype_code == F64 or y_is_fp64) else np.float32
    yd = DeviceArray.zeros(D * K, ydt)
    spined.matvec_rows_device(op, xd, yd, dtype_code, K, y_is_fp64=y_is_fp64)
    spined.synchronize()
    return yd.numpy().reshape(D, K)


def _single_matvec(op, D, x, dtype_code, y_is_fp64=False):
    xd = DeviceArray.from_numpy(np.ascontiguousarray(x))
    ydt = np.float64 if (dtype_code == F64 or y_is_fp64) else np.float32
    yd = DeviceArray.zeros(D, ydt)
    spined.matvec_device(op, xd, yd, dtype_code, y_is_fp64=y_is_fp64)
    spined.synchronize()
    return yd.numpy()


@pytest.mark.parametrize("spec", BLOCK_CASES, ids=lambda s: s.name)
@pytest.mark.parametrize("K", [2, 4, 8])
def test_block_matches_single_fp64(spec, K):
    print("case", spec.name, "K", K, flush=True)
    basis, op = engine(spec)
    o = oracle(spec)
    D = o.dimension
    X = np.random.default_rng(11 + K).uniform(-1, 1, (D, K))
    for method in ([0, 1] if spined.ring_available(basis) else [0]):
        spined.set_orbit_method(op, method)
        Y = _rows_matvec(op, D, X, F64)
        for c in range(K):
            ys = _single_matvec(op, D, X[:, c], F64)
            np.testing.assert_array_equal(Y[:, c], ys, err_msg=f"method {method} column {c}")
        yr = o.matvec(X[:, K - 1])
        assert np.max(np.abs(Y[:, K - 1] - yr)) <= 1e-13 * np.max(np.abs(yr))


@pytest.mark.parametrize("K", [4, 8])
def test_block_fp32_storage(K):
    spec = ring(22)
    basis, op = engine(spec)
    o = oracle(spec)
    D = o.dimension
    X = np.random.default_rng(7).uniform(-1, 1, (D, K)).astype(np.float32)
    Ym = _rows_matvec(op, D, X, F32, y_is_fp64=True)
    Y32 = _rows_matvec(op, D, X, F32)
    assert Y32.dtype == np.float32
    for c in range(K):
        np.testing.assert_array_equal(Ym[:, c], _single_matvec(op, D, X[:, c], F32, True))
        np.testing.assert_array_equal(Y32[:, c], _single_matvec(op, D, X[:, c], F32))
    yr = o.matvec(X[:, 0].astype(np.float64))
    assert np.max(np.abs(Ym[:, 0] - yr)) <= 1e-13 * np.max(np.abs(yr))


@pytest.mark.parametrize("block", [2, 5, 11])
def test_ls_operator_matmat_block(block):
    """Host ls_operator_matmat with block_size > 1 (chunks of 8/4/2/1 columns)
    equals the oracle column by column; strides larger than the size."""
    spec = ring(16, k=0, p=0, z=1)
    basis, op = engine(spec)
    o = oracle(spec)
    D = o.dimension
    stride = D + 3
    rng = np.random.default_rng(block)
    buf = rng.uniform(-1, 1, stride * block)
    X = np.lib.stride_tricks.as_strided(buf, shape=(D, block), strides=(8, 8 * stride))
    ybuf = np.full(stride * block, 7.0)
    from spined import _lib
    st = _lib.load().ls_operator_matmat(op.ptr, F64, D, block, buf.ctypes.data, stride,
                                        ybuf.ctypes.data, stride)
    assert st == 0
    for c in range(block):
        yr = o.matvec(np.ascontiguousarray(X[:, c]))
        yc = ybuf[c * stride:c * stride + D]
        assert np.max(np.abs(yc - yr)) <= 1e-13 * np.max(np.abs(yr)), c
        # padding between columns untouched
        assert np.all(ybuf[c * stride + D:(c + 1) * stride] == 7.0)


def test_rows_device_rejects_bad_width():
    spec = ring(12)
    basis, op = engine(spec)
    D = spined.get_number_states(basis)
    xd = DeviceArray.zeros(D * 3, np.float64)
    yd = DeviceArray.zeros(D * 3, np.float64)
    with pytest.raises(Exception):
        spined.matvec_rows_device(op, xd, yd, F64, 3)


# ---- block Lanczos phase (max_block_size >= 2) ---------------------------------
RING24_E0 = -42.680058066148944  # Bethe ansatz, N = 24 Heisenberg ring


@pytest.mark.parametrize("K", [2, 4, 8])
def test_block_eigsh_fp64_ring24(K):
    spec = ring(24)
    basis, op = engine(spec)
    r = spined.eigsh(op, k=1, tol=1e-10, max_block_size=K)
    print("K", K, r.evals, r.rnorms, r.stats, flush=True)
    assert r.status == 0
    assert abs(r.evals[0] - RING24_E0) <= 1e-10 * abs(RING24_E0)
    assert r.rnorms[0] <= 1e-10
    o = oracle(spec)
    ref = o.lanczos(k=1, tol=1e-11)
    assert abs(np.dot(r.evecs[:, 0], ref["evecs"][:, 0])) >= 1 - 1e-8


def test_block_eigsh_lowest4_ring24():
    spec = ring(24)
    basis, op = engine(spec)
    r = spined.eigsh(op, k=4, tol=1e-10, max_block_size=4)
    o = oracle(spec)
    ref = o.lanczos(k=4, tol=1e-11, max_iter=400)
    print(r.evals, r.rnorms, r.stats, flush=True)
    assert r.status == 0
    np.testing.assert_allclose(r.evals, ref["evals"], rtol=1e-10, atol=0)
    assert np.all(r.rnorms <= 1e-10)


def test_block_eigsh_mixed_ring26():
    spec = ring(26, k=13, p=1, z=-1)   # N/2 odd: ground state at momentum pi
    basis, op = engine(spec)
    r = spined.eigsh(op, k=1, tol=1e-10, precision="mixed", max_block_size=8)
    ref = spined.eigsh(op, k=1, tol=1e-10)
    print(r.evals, r.rnorms, r.stats, ref.evals, flush=True)
    assert r.status == 0 and ref.status == 0
    assert abs(r.evals[0] - ref.evals[0]) <= 1e-10 * abs(ref.evals[0])
    assert r.rnorms[0] <= 1e-10
    assert r.stats["matvecs_mixed"] > 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_block_eigsh_invariant_subspace(precision, capfd):
    """One exchange bond on 18 sites (w = 9, D = 48,620): H = 1 - 4 P (P the
    bond's singlet projector), so the second block of the Krylov space
    already closes it and the Gram-Schmidt of the next block cancels to
    rounding: the second pass runs there (the log counts it) and the block
    turns rank deficient; the solve still ends at E0 = -3."""
    basis = spined.mk_basis(spined.mk_group([]), 18, 9, None)
    op = spined.mk_operator(basis, [spined.mk_interaction(lattice.HEISENBERG, [[0, 1]])])
    spined.build_basis(basis)
    r = spined.eigsh(op, k=1, tol=1e-10, max_block_size=8, precision=precision, verbose=True,
                     max_matvecs=3000)
    err = capfd.readouterr().err
    print(r.evals, r.rnorms, r.stats, flush=True)
    assert r.status == 0
    assert abs(r.evals[0] + 3.0) <= 1e-12
    assert r.rnorms[0] <= 1e-10
    counts = [int(line.rsplit(" ", 1)[1]) for line in err.splitlines()
              if "second Gram-Schmidt passes" in line]
    assert counts and max(counts) >= 1


# ---- row schedule (rows of equal work side by side) ----------------------------
@pytest.mark.parametrize("spec", [ring(20, k=0, p=0, z=1), ring(24), ring(16, k=0, p=0, z=1),
                                  ring(12, k=0, p=0, z=None, w=None, matrix=SIGMA_XX,
                                       name="xx12_templates")],
                         ids=lambda s: s.name)
def test_schedule_bit_identical(spec):
    """The scheduled kernels compute every row exactly as the unscheduled
    ones: single (FP64 and FP32-storage) and block matvecs are bit-identical;
    forced on for bases smaller than one tile (a single partial tile)."""
    print("case", spec.name, flush=True)
    basis, op = engine(spec)
    D = spined.get_number_states(basis)
    rng = np.random.default_rng(17)
    x = rng.uniform(-1, 1, D)
    X = rng.uniform(-1, 1, (D, 4))
    outs = []
    for mode in (0, 1):
        spined.set_schedule(op, mode)
        y = spined.apply(op, x)
        assert spined.schedule(op) == mode
        ym = _single_matvec(op, D, x.astype(np.float32), F32, True)
        Y = _rows_matvec(op, D, X, F64)
        outs.append((y, ym, Y))
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a, b)
    yr = oracle(spec).matvec(x)
    assert np.max(np.abs(outs[1][0] - yr)) <= 1e-13 * np.max(np.abs(yr))
    spined.set_schedule(op, -1)
    spined.apply(op, x)
    assert spined.schedule(op) == (1 if D >= 4096 else 0)
