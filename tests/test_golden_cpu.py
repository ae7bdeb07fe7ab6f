"""The oracle against the committed golden fixtures (tests/golden, written
by tests/golden/make_golden.py from the example decks shipped with SpinED):
a regression guard on the checker itself.  Decks whose oracle build takes
more than a few seconds are left to the GPU tests."""
import numpy as np
import pytest

import pyoracle
from golden_util import NAMES, load, matrices

FAST = [n for n in NAMES if n not in ("heisenberg_square_5x5", "heisenberg_pyrochlore_32")]


@pytest.mark.parametrize("name", FAST)
def test_oracle_reproduces_fixture(name):
    g = load(name)
    deck = g["deck"]
    o = pyoracle.System(deck["number_spins"], deck["hamming_weight"], deck["spin_inversion"],
                        [(p, s) for p, s in deck["symmetries"]])
    for m, s in matrices(deck):
        o.add_interaction(m, s)
    o.build()
    reps, stab = o.basis()
    D = len(reps)
    assert D == int(g["dimension"])
    assert np.bitwise_xor.reduce(reps) == g["reps_xor"]
    assert np.sum(reps, dtype=np.uint64) == g["reps_sum"]
    np.testing.assert_array_equal(reps[g["sample_idx"]], g["sample_rep"])
    np.testing.assert_array_equal(stab[g["sample_idx"]], g["sample_stab"])
    x = pyoracle.random_vector(int(g["x_seeds"][0]), D)
    if not o.is_real:
        x = x + 1j * pyoracle.random_vector(int(g["x_seeds"][1]), D)
    y = o.matvec(x)
    np.testing.assert_array_equal(np.real(y[g["y_idx"]]), g["y_re"])
    np.testing.assert_array_equal(np.imag(y[g["y_idx"]]), g["y_im"])
