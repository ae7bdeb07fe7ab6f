"""Combinatorial-rank index (index kind 2) of bases whose symmetry group has
no permutation but the identity: index(v) = sum_i C(p_i, i + 1) over the up
sites p_0 < p_1 < ... (weight-w sectors, with or without the spin flip), or
index(v) = v without a weight.  It is chosen automatically for such bases.

Every lookup is exact in all three index forms, so the rank index must give
the same matvecs (single, block, complex, FP32), off-diagonal counts, gather
caches and eigenpairs as the bucket table and the hash table, bit for bit,
and the oracle's matvec to 1e-13.  Covered: all five sector shapes (weight
w, weight n/2 with either flip sector, all weights with and without the
flip), an operator that leaves the weight sector (its images are rejected by
their weight), 3-site terms, a complex operator, the one-state sectors w = 0
and w = n, and the refusal on bases the rank does not describe.
"""
import numpy as np
import pytest

import spined
from spined import lattice
from spined.device import DeviceArray
from spined.internal import LatticeSymmetriesException
from helpers import Spec, engine, oracle, ring
from test_oracle_pins import random_hermitian

pytestmark = pytest.mark.gpu

SX = [[0, 1], [1, 0]]


def xxz_chain(n, seed, periodic=True):
    rng = np.random.default_rng(seed)
    nb = n if periodic else n - 1
    return [(lattice.xxz_matrix(float(rng.uniform(0, 2)), float(rng.uniform(0.8, 1.2))),
             [[k, (k + 1) % n]]) for k in range(nb)]


def _cases():
    c = []
    c.append(Spec(16, 8, None, [], xxz_chain(16, 1, periodic=False) +
                  [(random_hermitian(3, 61, real=True, conserve_weight=True),
                    [[i, i + 1, i + 2] for i in range(14)])], "chain16_w8_xxz_3site"))
    c.append(Spec(18, 9, 1, [], xxz_chain(18, 2), "ring18_w9_flip_plus"))
    c.append(Spec(14, 7, -1, [], xxz_chain(14, 3), "ring14_w7_flip_minus"))
    c.append(Spec(12, None, None, [], xxz_chain(12, 4) + [(SX, [[i] for i in range(12)])],
                  "chain12_all_weights_field"))
    c.append(Spec(12, None, -1, [], xxz_chain(12, 5) + [(SX, [[i] for i in range(12)])],
                  "chain12_all_weights_flip_minus_field"))
    c.append(Spec(14, 5, None, [], xxz_chain(14, 6) + [(SX, [[i] for i in range(14)])],
                  "chain14_w5_field_leaves_sector"))
    c.append(Spec(12, 6, None, [], [(random_hermitian(2, 62, conserve_weight=True),
                                     [[k, (k + 1) % 12] for k in range(12)])],
                  "ring12_w6_complex"))
    c.append(Spec(8, 0, None, [], xxz_chain(8, 7), "ring8_w0_one_state"))
    c.append(Spec(8, 8, None, [], xxz_chain(8, 8), "ring8_w8_one_state"))
    return c


@pytest.mark.parametrize("spec", _cases(), ids=lambda s: s.name)
def test_rank_index_matches_bucket_and_hash(spec):
    print("case", spec.name, flush=True)
    basis, op = engine(spec)
    o = oracle(spec)
    D = o.dimension
    assert spined.get_number_states(basis) == D
    np.testing.assert_array_equal(spined.basis_get_states(basis), o.basis()[0])
    assert spined.index_kind(basis) == 2
    rng = np.random.default_rng(31)
    x = rng.uniform(-1, 1, D)
    if not o.is_real:
        x = x + 1j * rng.uniform(-1, 1, D)
    y2 = spined.apply(op, x)
    c2 = spined.count_offdiag(op)
    yr = o.matvec(x)
    assert np.max(np.abs(y2 - yr)) <= 1e-13 * max(np.max(np.abs(yr)), 1.0)
    for kind in (0, 1):
        spined.set_index_kind(basis, kind)
        assert spined.index_kind(basis) == kind
        np.testing.assert_array_equal(spined.apply(op, x), y2, err_msg=f"kind {kind}")
        assert spined.count_offdiag(op) == c2, kind
    spined.set_index_kind(basis, -1)
    assert spined.index_kind(basis) == 2
    np.testing.assert_array_equal(spined.apply(op, x), y2)
    if o.is_real:
        x32 = x.astype(np.float32)
        y32 = spined.apply(op, x32)
        spined.set_index_kind(basis, 0)
        np.testing.assert_array_equal(spined.apply(op, x32), y32)
        spined.set_index_kind(basis, -1)
        K = 4
        X = rng.uniform(-1, 1, (D, K))
        xd = DeviceArray.from_numpy(np.ascontiguousarray(X))
        yd = DeviceArray.zeros(D * K, np.float64)
        spined.matvec_rows_device(op, xd, yd, 1, K)
        spined.synchronize()
        Y = yd.numpy().reshape(D, K)
        for c in range(K):
            np.testing.assert_array_equal(Y[:, c], spined.apply(op, np.ascontiguousarray(X[:, c])))
        # the gather cache resolves its targets through the same lookups
        spined.set_gather_cache(op, 1)
        np.testing.assert_array_equal(spined.apply(op, x), y2)
        spined.set_gather_cache(op, 0)


def test_rank_index_eigsh_identical():
    spec = _cases()[1]
    basis, op = engine(spec)
    assert spined.index_kind(basis) == 2
    r2 = spined.eigsh(op, k=1, tol=1e-10)
    spined.set_index_kind(basis, 0)
    r0 = spined.eigsh(op, k=1, tol=1e-10)
    assert r2.status == 0 and r0.status == 0
    assert r2.evals[0] == r0.evals[0]
    e = oracle(spec).lanczos(k=1, tol=1e-11, max_iter=400)["evals"][0]
    assert abs(r2.evals[0] - e) <= 1e-10 * abs(e)


def test_rank_index_refused_where_it_does_not_apply():
    # translations: the representatives are orbit minima, not every state
    b, _ = engine(ring(12))
    assert spined.index_kind(b) == 0
    with pytest.raises(LatticeSymmetriesException):
        spined.set_index_kind(b, 2)
    assert spined.index_kind(b) == 0
    # a partial list of a symmetry-free sector is not complete: no rank
    spec = Spec(12, 6, None, [], xxz_chain(12, 9))
    full, op_full = engine(spec)
    assert spined.index_kind(full) == 2
    reps = spined.basis_get_states(full)
    keep = np.ones(len(reps), dtype=bool)
    keep[::5] = False
    part, op_part = engine(spec, build=False)
    spined.build_basis(part, reps[keep])
    assert spined.is_complete(part) == 0
    assert spined.index_kind(part) == 0
    with pytest.raises(LatticeSymmetriesException):
        spined.set_index_kind(part, 2)
    # the whole list adopted through ls_build_unsafe is complete: rank again
    adopted, op_ad = engine(spec, build=False)
    spined.build_basis(adopted, reps)
    assert spined.index_kind(adopted) == 2
    x = np.random.default_rng(3).uniform(-1, 1, len(reps))
    np.testing.assert_array_equal(spined.apply(op_ad, x), spined.apply(op_full, x))


def test_autotune_offers_rank_index():
    spec = _cases()[1]
    basis, op = engine(spec)
    D = spined.get_number_states(basis)
    x = np.random.default_rng(8).uniform(-1, 1, D)
    y0 = spined.apply(op, x)
    rep = spined.autotune(op, repeats=2)
    print(rep, flush=True)
    assert rep["index_kind"] in (0, 2)
    assert spined.index_kind(basis) == rep["index_kind"]
    np.testing.assert_array_equal(spined.apply(op, x), y0)
