"""Rows per wave of the matrix-free real matvec of exchange-bond operators
(no reference counterpart: the rows-per-wavefront setting of the
autotuner).  At 64 rows per wave every lane carries a row; at 32 or 16 a
row's active templates are spread over 2 or 4 lanes, and every lane of the
row accumulates the terms in template order with the same fma chain, so all
three settings give bit-identical sums."""
import numpy as np
import pytest

import spined
from spined import lattice
from helpers import Spec, engine, oracle, ring

pytestmark = pytest.mark.gpu


def reach_bonds(n, reach):
    # bonds (i, i + d mod n) for d = 1 .. reach: n * reach bonds
    return [[i, (i + d) % n] for d in range(1, reach + 1) for i in range(n)]


CASES = [
    ring(24, k=0, p=0, z=1),   # ring orbits of nearest-neighbour moves (hop path)
    Spec(16, 8, 1, lattice.square_symmetries(4), [(lattice.HEISENBERG, lattice.square_bonds(4))],
         "square4x4"),         # generic group programs
    Spec(18, 9, 1, [], [(lattice.xxz_matrix(0.7), lattice.ring_bonds(18))],
         "xxz18_rank"),        # no permutations: rank index and nearest-neighbour rank moves
    Spec(20, 10, 1, lattice.ring_symmetries(20, 0, 0), [(lattice.HEISENBERG, reach_bonds(20, 4))],
         "ring20_reach4"),     # 80 bonds: two 64-bit template words, full ring scan
]


@pytest.mark.parametrize("spec", CASES, ids=lambda s: s.name)
def test_rows_per_wave_bit_identical(spec):
    print("case", spec.name, flush=True)
    basis, op = engine(spec)
    o = oracle(spec)
    spined.set_gather_cache(op, 0)
    D = o.dimension
    rng = np.random.default_rng(17)
    x64 = rng.uniform(-1, 1, D)
    x32 = rng.uniform(-1, 1, D).astype(np.float32)
    yr = o.matvec(x64)
    for method in ([0, 1] if spined.ring_available(basis) else [0]):
        spined.set_orbit_method(op, method)
        ys = {}
        for rows in (64, 32, 16):
            spined.set_rows_per_wave(op, rows)
            assert spined.rows_per_wave(op) == rows
            ys[rows] = (spined.apply(op, x64), spined.apply(op, x32))
        for rows in (32, 16):
            assert np.array_equal(ys[rows][0], ys[64][0]), (method, rows)
            assert np.array_equal(ys[rows][1], ys[64][1]), (method, rows)
        assert np.max(np.abs(ys[16][0] - yr)) <= 1e-13 * np.max(np.abs(yr)), method
    spined.set_rows_per_wave(op, -1)
    assert spined.rows_per_wave(op) == 64


def test_rows_per_wave_solve_and_bad_setting():
    basis, op = engine(ring(20, k=0, p=0, z=1))
    spined.set_gather_cache(op, 0)
    r64 = spined.eigsh(op, k=1, tol=1e-10, precision="mixed")
    spined.set_rows_per_wave(op, 16)
    r16 = spined.eigsh(op, k=1, tol=1e-10, precision="mixed")
    assert r16.status == 0 and r16.rnorms[0] <= 1e-10
    # identical matvecs: the same recurrence, the same eigenvalue bits
    assert r16.evals[0] == r64.evals[0]
    assert r16.stats["matvecs"] == r64.stats["matvecs"]
    with pytest.raises(Exception):
        spined.set_rows_per_wave(op, 8)


def test_autotune_times_rows_per_wave():
    basis, op = engine(ring(22, k=0, p=0, z=1))
    spined.set_gather_cache(op, 0)
    x = np.random.default_rng(23).uniform(-1, 1, spined.get_number_states(basis))
    y0 = spined.apply(op, x)
    rep = spined.autotune(op, repeats=2)
    print(rep)
    assert rep["rows_per_wave"] in (64, 32, 16)
    assert set(rep["ms_per_rows_per_wave"]) == {64, 32, 16}
    assert spined.rows_per_wave(op) == rep["rows_per_wave"]
    assert np.array_equal(spined.apply(op, x), y0)
