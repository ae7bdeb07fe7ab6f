"""Hash index of representatives (spined_basis_set_index_kind(basis, 1))
against the bucket index (and, for the basis without permutations, the
rank index): every lookup is exact in each form, so matvecs
(single, block, complex), off-diagonal counts and eigenpairs are
bit-identical."""
import numpy as np
import pytest

import spined
from spined import lattice
from spined.device import DeviceArray
from helpers import Spec, engine, oracle, ring

pytestmark = pytest.mark.gpu

CASES = [
    ring(10, k=5, p=1, z=-1, name="chain10_example"),
    ring(20, k=0, p=0, z=1),
    ring(20, k=3, p=None, z=None, name="ring20_k3_complex"),
    ring(18, k=9, p=1, z=-1),
    ring(14, k=None, p=None, z=1, name="xxz14_flip"),
    Spec(16, 8, 1, lattice.square_symmetries(4), [(lattice.HEISENBERG, lattice.square_bonds(4))],
         "square4x4"),
]


@pytest.mark.parametrize("spec", CASES, ids=lambda s: s.name)
def test_hash_index_matvec_bit_identical(spec):
    print("case", spec.name, flush=True)
    basis, op = engine(spec)
    o = oracle(spec)
    D = o.dimension
    rng = np.random.default_rng(21)
    x = rng.uniform(-1, 1, D)
    if not o.is_real:
        x = x + 1j * rng.uniform(-1, 1, D)
    # a basis with no permutation symmetry gets the rank index by default
    assert spined.index_kind(basis) == (2 if not spec.gens else 0)
    y0 = spined.apply(op, x)
    c0 = spined.count_offdiag(op)
    spined.set_index_kind(basis, 1)
    assert spined.index_kind(basis) == 1
    y1 = spined.apply(op, x)
    c1 = spined.count_offdiag(op)
    np.testing.assert_array_equal(y0, y1)
    assert c0 == c1
    yr = o.matvec(x)
    assert np.max(np.abs(y1 - yr)) <= 1e-13 * np.max(np.abs(yr))
    if o.is_real:
        K = 4
        X = rng.uniform(-1, 1, (D, K))
        xd = DeviceArray.from_numpy(np.ascontiguousarray(X))
        yd = DeviceArray.zeros(D * K, np.float64)
        spined.matvec_rows_device(op, xd, yd, 1, K)
        spined.synchronize()
        Y = yd.numpy().reshape(D, K)
        for c in range(K):
            np.testing.assert_array_equal(Y[:, c], spined.apply(op, np.ascontiguousarray(X[:, c])))
    spined.set_index_kind(basis, 0)
    np.testing.assert_array_equal(spined.apply(op, x), y0)


def test_hash_index_before_build_and_eigsh():
    spec = ring(24)
    syms = [spined.mk_symmetry(p, s) for p, s in spec.gens]
    basis = spined.mk_basis(spined.mk_group(syms), spec.n, spec.w, spec.z)
    spined.set_index_kind(basis, 1)
    spined.build_basis(basis)
    assert spined.index_kind(basis) == 1
    op = spined.mk_operator(basis, [spined.mk_interaction(m, s) for m, s in spec.terms])
    r = spined.eigsh(op, k=1, tol=1e-10)
    assert r.status == 0
    assert abs(r.evals[0] - (-42.680058066148944)) <= 1e-10 * 42.68
    assert r.rnorms[0] <= 1e-10


@pytest.mark.parametrize("spec", CASES[:4], ids=lambda s: s.name)
def test_bucket_sizes_bit_identical(spec):
    """The two-level bucket index at very coarse, default and very fine level-2
    sizes: every lookup is exact, so matvecs and gather counts do not change."""
    print("case", spec.name, flush=True)
    basis, op = engine(spec)
    o = oracle(spec)
    D = o.dimension
    x = np.random.default_rng(5).uniform(-1, 1, D)
    if not o.is_real:
        x = x + 1j * np.random.default_rng(6).uniform(-1, 1, D)
    y0 = spined.apply(op, x)
    c0 = spined.count_offdiag(op)
    default_bits = spined.bucket_bits(basis)
    for bits in (1, 4, default_bits + 2, 36, -1):
        spined.set_bucket_bits(basis, bits)
        np.testing.assert_array_equal(spined.apply(op, x), y0, err_msg=f"bits {bits}")
        assert spined.count_offdiag(op) == c0, bits
    assert spined.bucket_bits(basis) == default_bits
    yr = o.matvec(x)
    assert np.max(np.abs(y0 - yr)) <= 1e-13 * np.max(np.abs(yr))


def test_autotune_keeps_results_exact():
    spec = ring(24)
    basis, op = engine(spec)
    D = spined.get_number_states(basis)
    x = np.random.default_rng(8).uniform(-1, 1, D)
    y0 = spined.apply(op, x)
    rep = spined.autotune(op, repeats=2)
    print(rep, flush=True)
    assert rep["candidates"] >= 4
    assert rep["ms_per_matvec"] > 0 and rep["ms_per_matvec_first"] > 0
    assert rep["best_block"] in (1, 2, 4, 8)
    assert set(rep["ms_per_vector_block"]) == {1, 2, 4, 8}
    assert spined.index_kind(basis) == rep["index_kind"]
    np.testing.assert_array_equal(spined.apply(op, x), y0)


SX = [[0, 1], [1, 0]]


def test_complete_flag_and_partial_lists():
    """The bucket search skips its confirming load of reps only on a basis
    that holds every valid orbit minimum of its sector (ls_build, or an
    ls_build_unsafe list that passes the check); a partial list keeps the
    load, so the images of the dropped states are not found."""
    spec = ring(16, k=0, p=1, z=1)
    b1, op1 = engine(spec)
    assert spined.is_complete(b1) == 1
    reps = spined.basis_get_states(b1)
    D = len(reps)
    x = np.random.default_rng(5).uniform(-1, 1, D)
    y1 = spined.apply(op1, x)
    b2, op2 = engine(spec, build=False)
    assert spined.is_complete(b2) == -1
    spined.build_basis(b2, reps)
    assert spined.is_complete(b2) == 1
    np.testing.assert_array_equal(spined.apply(op2, x), y1)
    keep = np.ones(D, dtype=bool)
    keep[::3] = False
    part = reps[keep]
    b3, op3 = engine(spec, build=False)
    spined.build_basis(b3, part)
    assert spined.is_complete(b3) == 0
    o = oracle(spec, build=False)
    o.set_basis(part)
    xp = x[keep]
    yo = o.matvec(xp)
    for bits in (1, 6, 36, -1):
        spined.set_bucket_bits(b3, bits)
        y3 = spined.apply(op3, xp)
        assert np.max(np.abs(y3 - yo)) <= 1e-13 * np.max(np.abs(yo)), bits
    # a list as long as the sector but holding a state that is not an orbit
    # minimum is not complete either
    bad = reps.copy()
    r0 = int(bad[0])
    alt = ((r0 << 1) | (r0 >> 15)) & 0xFFFF       # a rotation of the first minimum
    assert alt not in set(reps.tolist()) and o.orbit(alt)[0] == r0
    bad[0] = alt
    bad = np.sort(bad)
    b4, _ = engine(spec, build=False)
    spined.build_basis(b4, bad)
    assert spined.is_complete(b4) == 0


@pytest.mark.parametrize("w", ["half", None], ids=["weight8", "all_weights"])
def test_weight_changing_operator(w):
    """An operator with a transverse field leaves the Hamming-weight sector:
    its images must be dropped on a weight-restricted basis (lookups keep
    the confirming load) and found on a basis of all weights."""
    spec = ring(16, k=0, p=1, z=1, w=w, name=f"ring16_sx_{w}")
    spec.terms.append((SX, [[i] for i in range(16)]))
    basis, op = engine(spec)
    assert spined.is_complete(basis) == 1
    o = oracle(spec)
    x = np.random.default_rng(6).uniform(-1, 1, o.dimension)
    yo = o.matvec(x)
    y = spined.apply(op, x)
    assert np.max(np.abs(y - yo)) <= 1e-13 * np.max(np.abs(yo))
    # the hash index always confirms: same gathers, same sums
    c = spined.count_offdiag(op)
    spined.set_index_kind(basis, 1)
    np.testing.assert_array_equal(spined.apply(op, x), y)
    assert spined.count_offdiag(op) == c
