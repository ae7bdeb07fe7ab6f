"""Heisenberg ring N = 40 at full size (Hamming weight 20, momentum 0,
reflection parity +1, spin inversion +1: 861,725,794 representatives;
example/heisenberg_chain_40.yaml).

Size-independent checks, since the oracle cannot enumerate the
C(40, 20) = 1.4e11 candidates in seconds:
  * the dimension equals the character-formula count;
  * orbit-stabiliser over the whole list: sum_r |G| / |Stab r| = C(40, 20)
    (|G| = 160), from the stabiliser sizes the engine stores;
  * the list is strictly ascending (so it is a set, in the order the index
    and the HDF5 layout assume);
  * 3,000 sampled representatives are the oracle's orbit minima with the
    oracle's stabiliser sizes;
  * three windows of 4,000 rows of H x equal the oracle's rows (the oracle
    indexes the engine's representatives) to 1e-13 of the row scale;
  * the headline solve -- learned start (CNN trained on the engine's own
    ground states of rings N = 16..32) with FP32 storage -- reaches the
    Bethe-ansatz energy to 1e-10 relative with residual <= 1e-10.
"""
import math

import numpy as np
import pytest

import pyoracle
import spined
from spined import lattice, learned
from helpers import engine, oracle, ring

pytestmark = pytest.mark.gpu

N = 40
SPEC = ring(N, k=0, p=0, z=1, name="ring40")
D_EXPECTED = 861725794
E_BETHE = -70.986091153332861
_cache = {}


def _system():
    if "sys" not in _cache:
        _cache["sys"] = engine(SPEC)
    return _cache["sys"]


def _states():
    if "reps" not in _cache:
        basis, _ = _system()
        _cache["reps"] = spined.basis_get_states(basis)
        _cache["stab"] = spined.basis_get_stabilizers(basis)
    return _cache["reps"], _cache["stab"]


def test_ring40_dimension_orbit_stabiliser_and_order():
    basis, _ = _system()
    D = spined.get_number_states(basis)
    assert D == D_EXPECTED
    assert round(spined.sector_dimension(basis)) == D
    reps, stab = _states()
    assert len(reps) == D and len(stab) == D
    G = 4 * N
    assert np.all(G % stab == 0)
    total = 0
    for b in range(0, D, 1 << 27):
        total += int(np.sum(G // stab[b:b + (1 << 27)].astype(np.int64)))
    assert total == math.comb(N, N // 2) == 137846528820
    for b in range(0, D, 1 << 27):
        seg = reps[b:min(D, b + (1 << 27) + 1)]
        assert np.all(seg[1:] > seg[:-1]), b
    assert bin(int(reps[0])).count("1") == N // 2
    assert int(reps[-1]) < (1 << N)


def test_ring40_sampled_orbit_minima():
    reps, stab = _states()
    o = oracle(SPEC, build=False)
    rng = np.random.default_rng(40)
    picks = np.concatenate([rng.choice(len(reps), 2990, replace=False),
                            np.arange(5), np.arange(len(reps) - 5, len(reps))])
    for i in picks:
        r, _, cnt, ok = o.orbit(int(reps[i]))
        assert ok and r == reps[i], i
        assert cnt == stab[i], i


def test_ring40_sampled_rows_match_oracle():
    reps, stab = _states()
    _, op = _system()
    D = len(reps)
    o = oracle(SPEC, build=False)
    o.set_basis_with_stab(reps, stab)
    x = pyoracle.random_vector(4040, D)
    y = spined.apply(op, x)
    for b in (0, D // 2, D - 4000):
        yr = o.matvec(x, rows=(b, b + 4000))
        scale = np.max(np.abs(yr))
        assert np.max(np.abs(y[b:b + 4000] - yr)) <= 1e-13 * scale, b
    del x, y


def test_ring40_mixed_learned_solve_reaches_bethe():
    _cache.pop("reps", None)
    _cache.pop("stab", None)
    _, op = _system()
    data = [learned.ground_state(m) for m in range(16, 34, 2)]
    model = learned.train(data, steps=1200, batch=2048, lr=3e-3)
    r = spined.eigsh(op, k=1, tol=1e-10, precision="mixed", model=model,
                     sign_mask=learned.marshall_mask(N), want_vectors=False, seed=0x5EED)
    print(r.evals, r.rnorms, r.stats, flush=True)
    assert r.status == 0
    assert r.rnorms[0] <= 1e-10
    assert abs(r.evals[0] - E_BETHE) <= 1e-10 * abs(E_BETHE)
    assert r.stats["matvecs_mixed"] > 0
