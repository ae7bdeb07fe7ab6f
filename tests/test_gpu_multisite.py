"""3- and 4-site interactions (ls_create_interaction3/4, Internal.hs:266-270,
:313-322) on the GPU against the oracle, whose matvec with such terms is
pinned by the dense Kronecker projection (test_oracle_pins.py, MULTI).

The matrices are random Hermitian ones that no site permutation leaves
alone, so a wrong local-index bit order (the first listed site is the most
significant bit) changes every row; the sets of tuples are translation
invariant, so the terms commute with the ring's translations.  Checked:
the orbit-and-index matvec (real and complex sectors, with and without a
Hamming weight), the gather cache (real sectors; bit-identical to the orbit
kernel, for operators within its 256 off-diagonal templates), and the
ground energy of a real sector against the oracle.
"""
import numpy as np
import pytest

import spined
from helpers import Spec, engine, oracle
from spined import lattice
from test_oracle_pins import random_hermitian, ring_tuples

pytestmark = pytest.mark.gpu


def _cases():
    n = 16
    return [
        Spec(n, 8, None, lattice.ring_symmetries(n, 0, None),
             [(random_hermitian(3, 131, real=True, conserve_weight=True), ring_tuples(n, 3)),
              (random_hermitian(4, 141, real=True, conserve_weight=True), ring_tuples(n, 4)),
              (lattice.HEISENBERG, lattice.ring_bonds(n))], "ring16_w8_k0_3site_4site"),
        Spec(n, None, None, lattice.ring_symmetries(n, 0, None),
             [(random_hermitian(3, 132, real=True), ring_tuples(n, 3)),
              (random_hermitian(4, 142, real=True), ring_tuples(n, 4))], "ring16_k0_all_weights"),
        Spec(n, 8, None, lattice.ring_symmetries(n, 5, None),
             [(random_hermitian(3, 133, conserve_weight=True), ring_tuples(n, 3)),
              (random_hermitian(4, 143, conserve_weight=True), ring_tuples(n, 4))],
             "ring16_w8_k5_complex"),
        Spec(14, 7, None, [],
             [(random_hermitian(4, 144, real=True, conserve_weight=True), ring_tuples(14, 4))],
             "chain14_w7_no_symmetry_4site"),
    ]


def ring_exchange(seed):
    """c (P + P^T) + a random diagonal on 4 sites, P the cyclic shift of the
    local bits (b0 b1 b2 b3) -> (b3 b0 b1 b2) with b0 the most significant:
    26 off-diagonal entries, no site-exchange symmetry."""
    rng = np.random.default_rng(seed)
    P = np.zeros((16, 16))
    for a in range(16):
        b = [(a >> (3 - i)) & 1 for i in range(4)]
        c = [b[3], b[0], b[1], b[2]]
        P[sum(v << (3 - i) for i, v in enumerate(c)), a] = 1.0
    return rng.uniform(0.5, 1.5) * (P + P.T) + np.diag(rng.normal(size=16))


def _cache_cases():
    # operators within the gather cache's 256 off-diagonal templates
    n = 16
    return [
        Spec(n, 8, None, lattice.ring_symmetries(n, 0, None),
             [(random_hermitian(3, 151, real=True, conserve_weight=True), ring_tuples(n, 3))],
             "ring16_w8_k0_3site_192_templates"),
        Spec(n, 8, None, [], [(ring_exchange(152), [[i, i + 1, i + 2, i + 3] for i in range(9)])],
             "chain16_w8_4site_ring_exchange_234_templates"),
    ]


@pytest.mark.parametrize("spec", _cases() + _cache_cases(), ids=lambda s: s.name)
def test_multisite_matvec_matches_oracle(spec):
    print("case", spec.name, flush=True)
    basis, op = engine(spec)
    o = oracle(spec)
    D = o.dimension
    assert spined.get_number_states(basis) == D
    np.testing.assert_array_equal(spined.basis_get_states(basis), o.basis()[0])
    rng = np.random.default_rng(7)
    if spined.is_operator_real(op):
        x = rng.uniform(-1, 1, D)
        spined.set_gather_cache(op, 0)
        y0 = spined.apply(op, x)
        yr = o.matvec(x)
        assert np.max(np.abs(y0 - yr)) <= 1e-13 * np.max(np.abs(yr))
        spined.set_gather_cache(op, 1)
        y1 = spined.apply(op, x)
        # beyond 256 off-diagonal templates the cache is not built and the
        # orbit kernel serves the matvec; either way the sums are the same
        if spec.name.endswith("_templates"):
            assert spined.gather_cache(op) == 1
        np.testing.assert_array_equal(y1, y0)
    else:
        x = rng.uniform(-1, 1, D) + 1j * rng.uniform(-1, 1, D)
        y = spined.apply(op, x.astype(np.complex128))
        yr = o.matvec(x.astype(np.complex128))
        assert np.max(np.abs(y - yr)) <= 1e-13 * np.max(np.abs(yr))


def test_multisite_ground_energy_matches_oracle():
    spec = _cases()[0]
    _, op = engine(spec)
    o = oracle(spec)
    ref = o.lanczos(k=1, tol=1e-11, max_iter=400)
    r = spined.eigsh(op, k=1, tol=1e-10)
    assert r.status == 0 and r.rnorms[0] <= 1e-10
    e = ref["evals"][0]
    assert abs(r.evals[0] - e) <= 1e-10 * abs(e), (r.evals[0], e)
