"""Candidate pruning of the basis builder at its edges, bit-exact against
the oracle's enumeration of every weight-w state.  A site-transitive group
(all translations of a periodic lattice) restricts the candidates to the
states with the top site empty; a ring with all rotations to those with
site 0 occupied too (x = (y << 1) | 1).  Cases: the smallest weights (w = 0
and w = 1 take the transitive branch only, w = 2 the ring branch with single-spin
y), the largest (w = n - 1: one candidate), flip sectors, non-trivial
momenta (orbit minima of any character), a translation subgroup that is not
transitive (no pruning) and a 2D lattice."""
import numpy as np
import pytest

import spined
from spined import lattice
from helpers import Spec, engine, oracle

pytestmark = pytest.mark.gpu


def _ring(n, w, k=0, p=0, z=None, name=None):
    return Spec(n, w, z, lattice.ring_symmetries(n, k, p),
                [(lattice.HEISENBERG, lattice.ring_bonds(n))], name or f"ring{n}_w{w}_k{k}_p{p}_z{z}")


CASES = [
    _ring(12, 0), _ring(12, 12), _ring(12, 1), _ring(12, 2), _ring(12, 3, k=1, p=None), _ring(12, 11), _ring(12, 10, k=3, p=None),
    _ring(14, 7, k=0, p=1, z=-1), _ring(13, 6, k=2, p=None), _ring(16, 8, z=1),
    Spec(12, 6, None, [([(i + 2) % 12 for i in range(12)], 0)],
         [(lattice.HEISENBERG, lattice.ring_bonds(12))], "ring12_w6_translation_by_2"),
    Spec(16, 5, None, lattice.square_symmetries(4), [(lattice.HEISENBERG, lattice.square_bonds(4))],
         "square4x4_w5"),
    Spec(16, 15, None, lattice.square_symmetries(4), [(lattice.HEISENBERG, lattice.square_bonds(4))],
         "square4x4_w15"),
]


@pytest.mark.parametrize("spec", CASES, ids=lambda s: s.name)
def test_pruned_basis_bit_exact(spec):
    print("case", spec.name, flush=True)
    basis, _ = engine(spec)
    o = oracle(spec)
    oreps, ostab = o.basis()
    reps = spined.basis_get_states(basis)
    np.testing.assert_array_equal(reps, oreps)
    np.testing.assert_array_equal(spined.basis_get_stabilizers(basis), ostab)
    d, _ = o.sector_dimension()
    assert len(reps) == round(d)
