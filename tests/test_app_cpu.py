"""Host-side front end and output layer, no GPU: the YAML parser mirrors the
FromJSON instances of src/SpinED.hs:81-173 (keys, optionality, defaults,
complex numbers as a number or [re, im], case-insensitive datatype, error
on a missing key), and the HDF5 layer writes and reads back the datasets of
SpinED's output (src/SpinED.hs:284-368) including complex ones as the
{r, i} compound."""
import numpy as np
import pytest

from spined import config, hdf5
from spined.config import ConfigError

DECK = """
basis:
  number_spins: 10
  hamming_weight: 5
  spin_inversion: -1
  symmetries:
    - permutation: [1, 2, 3, 4, 5, 6, 7, 8, 9, 0]
      sector: 5
    - permutation: [9, 8, 7, 6, 5, 4, 3, 2, 1, 0]
      sector: 1
hamiltonian:
  name: "Heisenberg Hamiltonian"
  terms:
    - matrix: [[1,  0,  0, 0],
               [0, -1,  2, 0],
               [0,  2, -1, 0],
               [0,  0,  0, 1]]
      sites: [[0, 1], [1, 2], [2, 3], [3, 4], [4, 5], [5, 6], [6, 7], [7, 8], [8, 9], [9, 0]]
observables:
  - name: "SzSz"
    terms:
      - matrix: [[1, 0, 0, 0], [0, -1, 0, 0], [0, 0, -1, 0], [0, 0, 0, 1]]
        sites: [[0, 1]]
  - name: "complex"
    terms:
      - matrix: [[0, [0, -1]], [[0, 1], 0]]
        sites: [[0]]
number_vectors: 2
precision: 1e-9
datatype: Float32
"""


def test_parse_full_deck():
    c = config.loads(DECK)
    assert c.basis.number_spins == 10 and c.basis.hamming_weight == 5
    assert c.basis.spin_inversion == -1
    assert [s.sector for s in c.basis.symmetries] == [5, 1]
    assert c.basis.symmetries[0].permutation == [1, 2, 3, 4, 5, 6, 7, 8, 9, 0]
    assert c.hamiltonian.name == "Heisenberg Hamiltonian"
    m = np.array(c.hamiltonian.terms[0].matrix)
    np.testing.assert_array_equal(m.real, [[1, 0, 0, 0], [0, -1, 2, 0], [0, 2, -1, 0], [0, 0, 0, 1]])
    assert len(c.hamiltonian.terms[0].sites) == 10
    assert [o.name for o in c.observables] == ["SzSz", "complex"]
    assert c.observables[1].terms[0].matrix[0][1] == complex(0, -1)
    assert c.observables[1].terms[0].matrix[1][0] == complex(0, 1)
    assert c.number_vectors == 2 and c.precision == 1e-9
    assert c.datatype == "float32"                     # case-insensitive
    # defaults (SpinED.hs:161-173 and the engine's own keys)
    assert c.output == "exact_diagonalization_result.h5"
    assert c.max_primme_basis_size == 0 and c.min_primme_restart_size == 0
    assert c.solver_precision == "fp64" and c.start_vector == "random"


def test_optional_basis_keys_and_engine_keys():
    c = config.loads("""
basis: {number_spins: 4, symmetries: []}
hamiltonian: {name: H, terms: []}
observables: []
output: out/x.h5
solver_precision: MIXED
start_vector: learned
model: params.npy
seed: 7
""")
    assert c.basis.hamming_weight is None and c.basis.spin_inversion is None
    assert c.output == "out/x.h5" and c.solver_precision == "mixed"
    assert c.start_vector == "learned" and c.model == "params.npy" and c.seed == 7


@pytest.mark.parametrize("text,msg", [
    ("hamiltonian: {name: H, terms: []}\nobservables: []", 'key "basis" not found'),
    ("basis: {symmetries: []}\nhamiltonian: {name: H, terms: []}\nobservables: []",
     'key "number_spins" not found'),
    ("basis: {number_spins: 2, symmetries: []}\nhamiltonian: {terms: []}\nobservables: []",
     'key "name" not found'),
    ("basis: {number_spins: 2, symmetries: []}\nhamiltonian: {name: H, terms: []}\n"
     "observables: []\ndatatype: float16", "expected either 'float32' or 'float64'"),
    ("basis: {number_spins: 2, symmetries: [{permutation: [1, 0]}]}\n"
     "hamiltonian: {name: H, terms: []}\nobservables: []", 'key "sector" not found'),
    ("basis: {number_spins: 2, symmetries: []}\nhamiltonian: {name: H, terms: "
     "[{matrix: [[1, [0, 1, 2]]], sites: [[0]]}]}\nobservables: []", "complex number"),
    ("basis: {number_spins: 2, symmetries: []}\nhamiltonian: {name: H, terms: []}\n"
     "observables: []\nsolver_precision: fp16", "solver_precision"),
])
def test_config_errors(text, msg):
    with pytest.raises(ConfigError) as e:
        config.loads(text)
    assert msg in str(e.value)


@pytest.mark.skipif(not hdf5.available(), reason="libhdf5 not found")
def test_hdf5_round_trip(tmp_path):
    p = tmp_path / "r.h5"
    rng = np.random.default_rng(0)
    reps = np.array([3, 5, 6, 9, 2 ** 63 + 1], dtype=np.uint64)
    ev64 = rng.normal(size=3)
    ev32 = rng.normal(size=3).astype(np.float32)
    vec = rng.normal(size=(2, 7))
    c128 = rng.normal(size=(2, 5)) + 1j * rng.normal(size=(2, 5))
    c64 = (rng.normal(size=4) + 1j * rng.normal(size=4)).astype(np.complex64)
    with hdf5.File(p, "a") as f:
        for g in ("/basis", "/hamiltonian", "/observables", "/_workspace"):
            f.create_group(g)
        f.create_group("/basis")                      # idempotent
        f.write("/basis/representatives", reps)
        f.write("/hamiltonian/eigenvalues", ev64)
        f.write("/hamiltonian/residuals", ev32)
        f.write("/hamiltonian/eigenvectors", vec)
        f.write("/observables/c", c128)
        f.write("/observables/c64", c64)
        f.write("/hamiltonian/eigenvalues", ev64 * 2)  # overwrite with a warning
        # nested groups made level by level; int64; a silent overwrite
        f.write("/_workspace/lanczos/matvecs", np.array([7], np.int64))
        f.write("/_workspace/lanczos/matvecs", np.array([-3, 2 ** 40], np.int64), warn=False)
        # unlink a dataset; a missing one is no error
        f.write("/_workspace/lanczos/start_vector", vec[0])
        f.delete("/_workspace/lanczos/start_vector")
        f.delete("/_workspace/lanczos/absent")
        assert not f.exists("/_workspace/lanczos/start_vector")
    with hdf5.File(p, "r") as f:
        assert f.exists("/basis/representatives") and not f.exists("/basis/missing")
        np.testing.assert_array_equal(f.read("/basis/representatives", np.uint64), reps)
        np.testing.assert_array_equal(f.read("/hamiltonian/eigenvalues", np.float64), ev64 * 2)
        np.testing.assert_array_equal(f.read("/hamiltonian/residuals", np.float32), ev32)
        np.testing.assert_array_equal(f.read("/observables/c", np.complex128), c128)
        np.testing.assert_array_equal(f.read("/_workspace/lanczos/matvecs", np.int64),
                                      np.array([-3, 2 ** 40], np.int64))
        np.testing.assert_array_equal(f.read("/observables/c64", np.complex64), c64)
        # (k, D) on disk, as SpinED writes the column-major block
        np.testing.assert_array_equal(hdf5.read_eigenvectors(f, 7), vec.T)
    # README's (D, k) layout is accepted too
    with hdf5.File(tmp_path / "t.h5", "a") as f:
        f.write("/hamiltonian/eigenvectors", vec.T.copy())
        np.testing.assert_array_equal(hdf5.read_eigenvectors(f, 7), vec.T)
