"""The engine against the golden fixtures (tests/golden; inputs are the
example decks shipped with SpinED, outputs from the pinned oracle):
representatives bit-exact (count, checksums, sampled rows with their
stabilisers), H x on sampled rows within 1e-13 of max |y| and its sums,
lowest eigenvalues within 1e-10 (complex sectors through the complex
Hermitian Lanczos)."""
import numpy as np
import pytest

import pyoracle
import spined
from golden_util import NAMES, load, matrices

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", NAMES)
def test_engine_matches_fixture(name):
    print("case", name, flush=True)
    g = load(name)
    deck = g["deck"]
    syms = [spined.mk_symmetry(p, s) for p, s in deck["symmetries"]]
    basis = spined.mk_basis(spined.mk_group(syms), deck["number_spins"], deck["hamming_weight"],
                            deck["spin_inversion"])
    spined.build_basis(basis)
    op = spined.mk_operator(basis, [spined.mk_interaction(m, s) for m, s in matrices(deck)])
    D = spined.get_number_states(basis)
    assert D == int(g["dimension"])
    reps = spined.basis_get_states(basis)
    stab = spined.basis_get_stabilizers(basis)
    assert np.bitwise_xor.reduce(reps) == g["reps_xor"]
    assert np.sum(reps, dtype=np.uint64) == g["reps_sum"]
    np.testing.assert_array_equal(reps[g["sample_idx"]], g["sample_rep"])
    np.testing.assert_array_equal(stab[g["sample_idx"]], g["sample_stab"])
    x = pyoracle.random_vector(int(g["x_seeds"][0]), D)
    if not op.is_real:
        x = x + 1j * pyoracle.random_vector(int(g["x_seeds"][1]), D)
    y = spined.apply(op, x)
    yr = g["y_re"] + 1j * g["y_im"]
    scale = max(np.max(np.abs(yr)), 1.0)
    assert np.max(np.abs(y[g["y_idx"]] - yr)) <= 1e-13 * scale
    assert abs(np.sum(np.real(y)) - g["y_sum_re"]) <= 1e-9 * max(1.0, np.sqrt(g["y_norm2"]) * np.sqrt(D))
    assert abs(np.vdot(y, y).real - g["y_norm2"]) <= 1e-11 * g["y_norm2"]
    k = len(g["evals"])
    if k:
        r = spined.eigsh(op, k=k, tol=1e-10, want_vectors=False)
        print(r.evals, g["evals"], r.rnorms, flush=True)
        assert r.status == 0
        np.testing.assert_allclose(r.evals, g["evals"], rtol=1e-10, atol=1e-10)
