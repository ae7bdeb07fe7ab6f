"""Gather cache (spined_operator_set_gather_cache): the target of every term,
written once by the orbit and index kernels, read by later real matvecs.
The sums must equal those of the orbit-and-index kernels bit for bit (same
terms, coefficients and order), for every kind of basis and operator the
cache covers: ring and generic groups, negative characters, rows and
targets with stabilisers larger than one (the ratio bit), the hash index, a
partial representative list, a weight-changing operator, FP32 x."""
import numpy as np
import pytest

import spined
from spined import lattice
from helpers import Spec, engine, oracle, ring, torus

pytestmark = pytest.mark.gpu

SX = [[0, 1], [1, 0]]


def _cases():
    sx = ring(14, k=0, p=0, z=1, w=None, name="ring14_sx_all_weights")
    sx.terms.append((SX, [[i] for i in range(14)]))
    return [
        ring(20, k=0, p=0, z=1),
        ring(20, k=0, p=1, z=-1, name="ring20_p-_z-"),
        ring(18, k=9, p=0, z=1, name="ring18_k9"),
        ring(16, k=0, p=0, z=1, name="ring16_small_orbits"),
        Spec(16, 8, 1, lattice.square_symmetries(4), [(lattice.HEISENBERG, lattice.square_bonds(4))],
             "square4x4"),
        ring(14, k=None, p=None, z=1, name="xxz14_flip"),
        sx,
        # generic orbits: torus factorisation with trivial characters (the
        # fill's minimum-only scan) and with phases; a group that is no torus
        torus(6, 4, reflect=True, z=1),
        torus(4, 4, kx=2, ky=0, z=1),
        Spec(16, 8, None, [([(i + 2) % 16 for i in range(16)], 0)],
             [(lattice.HEISENBERG, lattice.ring_bonds(16))], "ring16_translation_by_2"),
    ]


def _check_same(op, D, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, D)
    spined.set_gather_cache(op, 0)
    y0 = spined.apply(op, x)
    y0_32 = spined.apply(op, x.astype(np.float32))
    spined.set_gather_cache(op, 1)
    y1 = spined.apply(op, x)
    assert spined.gather_cache(op) == 1
    np.testing.assert_array_equal(y1, y0)
    np.testing.assert_array_equal(spined.apply(op, x.astype(np.float32)), y0_32)
    return x, y1


@pytest.mark.parametrize("spec", _cases(), ids=lambda s: s.name)
def test_cache_matvec_bit_identical(spec):
    print("case", spec.name, flush=True)
    basis, op = engine(spec)
    o = oracle(spec)
    D = o.dimension
    x, y = _check_same(op, D, 3)
    yr = o.matvec(x)
    assert np.max(np.abs(y - yr)) <= 1e-13 * np.max(np.abs(yr))
    # a different row schedule drops the cache; rebuilt at the next matvec
    spined.set_schedule(op, 0)
    assert spined.gather_cache(op) == 0
    np.testing.assert_array_equal(spined.apply(op, x), y)
    assert spined.gather_cache(op) == 1
    spined.set_schedule(op, -1)
    # hash index: same targets
    spined.set_index_kind(basis, 1)
    spined.set_gather_cache(op, 0)
    spined.set_gather_cache(op, 1)
    np.testing.assert_array_equal(spined.apply(op, x), y)


def test_cache_partial_list():
    spec = ring(16, k=0, p=0, z=1)
    b1, _ = engine(spec)
    reps = spined.basis_get_states(b1)
    keep = np.ones(len(reps), dtype=bool)
    keep[::5] = False
    b2, op2 = engine(spec, build=False)
    spined.build_basis(b2, reps[keep])
    assert spined.is_complete(b2) == 0
    _check_same(op2, int(keep.sum()), 4)


def test_cache_eigsh_identical_and_auto():
    spec = ring(28, k=0, p=0, z=1)
    basis, op = engine(spec)
    D = spined.get_number_states(basis)
    res = {}
    for mode in (0, 1):
        spined.set_gather_cache(op, mode)
        for prec in ("fp64", "mixed"):
            r = spined.eigsh(op, k=1, tol=1e-10, precision=prec, want_vectors=False)
            assert r.status == 0 and r.rnorms[0] <= 1e-10
            res[(mode, prec)] = (r.evals[0], r.rnorms[0], r.stats["matvecs"])
    assert res[(0, "fp64")] == res[(1, "fp64")]
    assert res[(0, "mixed")] == res[(1, "mixed")]
    # automatic: bases of 2^20 rows and more, built at the first matvec
    big = ring(34, k=0, p=0, z=1)
    b3, op3 = engine(big)
    assert spined.get_number_states(b3) >= 1 << 20
    assert spined.gather_cache(op3) == 0
    spined.set_gather_cache(op3, -1)
    x = np.random.default_rng(5).uniform(-1, 1, spined.get_number_states(b3))
    y = spined.apply(op3, x)
    assert spined.gather_cache(op3) == 1
    spined.set_gather_cache(op3, 0)
    np.testing.assert_array_equal(spined.apply(op3, x), y)


@pytest.mark.parametrize("K", [2, 4, 8])
def test_cache_block_matvec_bit_identical(K):
    """Row-interleaved blocks (spined_operator_matmat_rows_device) from the
    gather cache: per column the sums of the orbit-and-index block kernel."""
    from spined.device import DeviceArray
    for spec in (ring(20, k=0, p=0, z=1),
                 Spec(16, 8, 1, lattice.square_symmetries(4),
                      [(lattice.HEISENBERG, lattice.square_bonds(4))], "square4x4")):
        print("case", spec.name, K, flush=True)
        basis, op = engine(spec)
        D = spined.get_number_states(basis)
        X = np.random.default_rng(K).uniform(-1, 1, (D, K))
        out = {}
        for mode in (0, 1):
            spined.set_gather_cache(op, mode)
            for code, dt in ((1, np.float64), (0, np.float32)):
                xd = DeviceArray.from_numpy(np.ascontiguousarray(X.astype(dt)))
                yd = DeviceArray.zeros(D * K, np.float64)
                spined.matvec_rows_device(op, xd, yd, code, K, y_is_fp64=code == 0)
                spined.synchronize()
                out[(mode, code)] = yd.numpy()
            assert spined.gather_cache(op) == mode
        np.testing.assert_array_equal(out[(1, 1)], out[(0, 1)])
        np.testing.assert_array_equal(out[(1, 0)], out[(0, 0)])
