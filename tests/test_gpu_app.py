"""The SpinED application flow end to end on the GPU (app/Main.hs runApp:
readConfig -> buildRepresentatives -> diagonalize -> computeExpectations ->
HDF5), checked against the oracle: the /basis/representatives dataset
bit-exact, eigenvalues within 1e-10, the (k, D) eigenvector block with
residuals within tolerance, observables as <v|O|v>; a second run on the
same output file takes the stored representatives (ls_build_unsafe, the
resume path of SpinED.hs:315-337); a complex momentum sector writes complex
eigenvectors."""
import numpy as np
import pytest

import pyoracle
import spined
from spined import app, hdf5, lattice

pytestmark = pytest.mark.gpu

RING10 = """
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
    - matrix: [[1, 0, 0, 0], [0, -1, 2, 0], [0, 2, -1, 0], [0, 0, 0, 1]]
      sites: [[0, 1], [1, 2], [2, 3], [3, 4], [4, 5], [5, 6], [6, 7], [7, 8], [8, 9], [9, 0]]
observables:
  - name: "SzSz"
    terms:
      - matrix: [[1, 0, 0, 0], [0, -1, 0, 0], [0, 0, -1, 0], [0, 0, 0, 1]]
        sites: [[0, 1]]
  - name: "SyOnSite0"
    terms:
      - matrix: [[0, [0, -1]], [[0, 1], 0]]
        sites: [[0]]
number_vectors: 2
precision: 1e-10
output: OUT
"""


def _oracle_ring10():
    gens = [([1, 2, 3, 4, 5, 6, 7, 8, 9, 0], 5), ([9, 8, 7, 6, 5, 4, 3, 2, 1, 0], 1)]
    o = pyoracle.System(10, 5, -1, gens)
    o.add_interaction(np.asarray(lattice.HEISENBERG, dtype=np.complex128),
                      lattice.ring_bonds(10))
    o.build()
    return o, gens


def _dense(o):
    D = o.dimension
    H = np.empty((D, D), dtype=np.complex128)
    for j in range(D):
        e = np.zeros(D, dtype=np.complex128)
        e[j] = 1.0
        H[:, j] = o.matvec(e)
    return H


@pytest.mark.skipif(not hdf5.available(), reason="libhdf5 not found")
def test_app_ring10_and_resume(tmp_path):
    out = tmp_path / "ring10.h5"
    deck = tmp_path / "ring10.yaml"
    deck.write_text(RING10.replace("OUT", str(out)))
    r = app.run_app(str(deck))
    o, gens = _oracle_ring10()
    reps, _ = o.basis()
    D = len(reps)
    ref = np.linalg.eigvalsh(_dense(o))[:2]
    with hdf5.File(out, "r") as f:
        np.testing.assert_array_equal(f.read("/basis/representatives", np.uint64), reps)
        ev = f.read("/hamiltonian/eigenvalues", np.float64)
        rn = f.read("/hamiltonian/residuals", np.float64)
        vecs = f.read("/hamiltonian/eigenvectors", np.float64)
        szsz = f.read("/observables/SzSz", np.complex128)
        sy = f.read("/observables/SyOnSite0", np.complex128)
    assert vecs.shape == (2, D)                        # (k, D): the column-major block
    np.testing.assert_allclose(ev, ref, rtol=1e-10, atol=1e-12)
    assert np.all(rn <= 1e-10)
    # observables: <v|O|v> for each eigenvector, against the oracle's operator
    ob = pyoracle.System(10, 5, -1, gens)
    ob.add_interaction(np.diag([1.0, -1.0, -1.0, 1.0]).astype(np.complex128), [[0, 1]])
    ob.set_basis(reps)
    for i in range(2):
        v = vecs[i]
        assert abs(np.linalg.norm(v) - 1) <= 1e-10
        assert np.linalg.norm(o.matvec(v) - ev[i] * v) <= 1e-9
        np.testing.assert_allclose(szsz[i].real, np.dot(v, ob.matvec(v)), rtol=1e-10, atol=1e-12)
    # sigma^y on one site has zero expectation in a real state
    assert np.max(np.abs(sy)) <= 1e-12
    # second run: representatives come from the file (resume path)
    r2 = app.run_app(str(deck))
    np.testing.assert_allclose(r2["evals"], r["evals"], rtol=1e-10)
    assert r2["dimension"] == D


@pytest.mark.skipif(not hdf5.available(), reason="libhdf5 not found")
def test_app_complex_sector(tmp_path):
    out = tmp_path / "ring12k5.h5"
    deck = tmp_path / "ring12k5.yaml"
    perm = [(i + 1) % 12 for i in range(12)]
    deck.write_text(f"""
basis:
  number_spins: 12
  hamming_weight: 6
  symmetries:
    - permutation: {perm}
      sector: 5
hamiltonian:
  name: H
  terms:
    - matrix: [[1, 0, 0, 0], [0, -1, 2, 0], [0, 2, -1, 0], [0, 0, 0, 1]]
      sites: {[[i, (i + 1) % 12] for i in range(12)]}
observables: []
number_vectors: 1
output: {out}
""")
    app.run_app(str(deck))
    o = pyoracle.System(12, 6, None, [(perm, 5)])
    o.add_interaction(np.asarray(lattice.HEISENBERG, dtype=np.complex128), lattice.ring_bonds(12))
    o.build()
    assert not o.is_real
    ref = np.linalg.eigvalsh(_dense(o))[0]
    with hdf5.File(out, "r") as f:
        ev = f.read("/hamiltonian/eigenvalues", np.float64)
        v = f.read("/hamiltonian/eigenvectors", np.complex128)
    assert abs(ev[0] - ref) <= 1e-10 * abs(ref)
    v = v.reshape(-1)
    assert np.linalg.norm(o.matvec(v) - ev[0] * v) <= 1e-9


def test_solver_checkpoints_and_resume():
    """spined_eigsh's checkpoint: after every restart the lowest Ritz
    vector of the current basis, unit norm, whose Rayleigh quotient is the
    reported Ritz value; a solve started from a mid-run checkpoint reaches
    the same E0 with fewer matvecs than the full run.  In the mixed solver
    the FP32-storage phase hands over float32 vectors, the FP64 phase
    float64 ones."""
    from helpers import engine, ring
    spec = ring(24)
    _, op = engine(spec)
    seen = []
    full = spined.eigsh(op, k=1, tol=1e-10, checkpoint=lambda v, info: seen.append((v, info)),
                        checkpoint_every=1)
    assert full.status == 0 and len(seen) >= 2
    assert [info["restarts"] for _, info in seen] == list(range(1, len(seen) + 1))
    for v, info in seen:
        assert v.dtype == np.float64
        assert abs(np.linalg.norm(v) - 1.0) <= 1e-8
        rq = float(np.dot(v, spined.apply(op, v)))
        assert abs(rq - info["ritz_values"][0]) <= 1e-8 * abs(rq)
    v, info = seen[len(seen) // 2]
    res = spined.eigsh(op, k=1, tol=1e-10, start_vector=v)
    assert res.status == 0
    assert abs(res.evals[0] - full.evals[0]) <= 1e-10 * abs(full.evals[0])
    assert res.stats["matvecs"] < full.stats["matvecs"]
    mixed = []
    r = spined.eigsh(op, k=1, tol=1e-10, precision="mixed",
                     checkpoint=lambda v, info: mixed.append((v.dtype, info["phase"])),
                     checkpoint_every=1)
    assert r.status == 0 and any(ph == 1 for _, ph in mixed)
    # FP32-storage phase: float32; FP64 phase (if it restarts at all): float64
    assert all((dt == np.dtype(np.float32)) == (ph == 1) for dt, ph in mixed)


@pytest.mark.skipif(not hdf5.available(), reason="libhdf5 not found")
def test_app_workspace_checkpoint_resume(tmp_path):
    """The application writes its Lanczos checkpoint to /_workspace/lanczos;
    at the end it marks the run converged there instead of storing the
    eigenvector a second time, and a second run on the same file starts from
    /hamiltonian/eigenvectors and needs only a few matvecs for the same E0.
    A run with a learned start evaluates the network on the device."""
    out = tmp_path / "ring24.h5"
    deck = tmp_path / "ring24.yaml"
    perm = [(i + 1) % 24 for i in range(24)]
    refl = [23 - i for i in range(24)]
    deck.write_text(f"""
basis:
  number_spins: 24
  hamming_weight: 12
  spin_inversion: 1
  symmetries:
    - permutation: {perm}
      sector: 0
    - permutation: {refl}
      sector: 0
hamiltonian:
  name: H
  terms:
    - matrix: [[1, 0, 0, 0], [0, -1, 2, 0], [0, 2, -1, 0], [0, 0, 0, 1]]
      sites: {[[i, (i + 1) % 24] for i in range(24)]}
observables: []
number_vectors: 1
precision: 1e-10
output: {out}
""")
    r = app.run_app(str(deck), checkpoint_every=2)
    assert not r["resumed"]
    D = r["dimension"]
    with hdf5.File(out, "r") as f:
        assert not f.exists("/_workspace/lanczos/start_vector")
        assert f.read("/_workspace/lanczos/converged", np.int64)[0] == 1
        rv = f.read("/_workspace/lanczos/ritz_values", np.float64)
        mv = f.read("/_workspace/lanczos/matvecs", np.int64)
        v = hdf5.read_eigenvectors(f, D)[:, 0]
    assert v.shape == (D,) and abs(np.linalg.norm(v) - 1.0) <= 1e-8
    np.testing.assert_allclose(rv[0], r["evals"][0], rtol=1e-12)
    assert mv[0] == r["matvecs"]
    r2 = app.run_app(str(deck), checkpoint_every=2)
    assert r2["resumed"]
    assert r2["matvecs"] < r["matvecs"] // 2
    np.testing.assert_allclose(r2["evals"], r["evals"], rtol=1e-10)
    assert abs(r["evals"][0] - (-42.680058066148944)) <= 1e-10 * 42.68   # Bethe ansatz, N = 24


@pytest.mark.skipif(not hdf5.available(), reason="libhdf5 not found")
def test_app_learned_start_on_device(tmp_path):
    """`start_vector: learned` with `model:` (a saved parameter file): the
    solver evaluates the network on the device and the mixed solve reaches
    the Bethe-ansatz energy (any network gives a valid start)."""
    from spined import learned
    from test_oracle_pins import bethe_ground_energy
    mp = tmp_path / "cnn.npy"
    learned.save(learned.CNN(seed=3), str(mp))
    out = tmp_path / "ring20.h5"
    deck = tmp_path / "ring20.yaml"
    n = 20
    deck.write_text(f"""
basis:
  number_spins: {n}
  hamming_weight: {n // 2}
  spin_inversion: 1
  symmetries:
    - permutation: {[(i + 1) % n for i in range(n)]}
      sector: 0
    - permutation: {[n - 1 - i for i in range(n)]}
      sector: 0
hamiltonian:
  name: H
  terms:
    - matrix: [[1, 0, 0, 0], [0, -1, 2, 0], [0, 2, -1, 0], [0, 0, 0, 1]]
      sites: {[[i, (i + 1) % n] for i in range(n)]}
observables: []
number_vectors: 1
precision: 1e-10
output: {out}
solver_precision: mixed
start_vector: learned
model: {mp}
""")
    r = app.run_app(str(deck), checkpoint_every=0)
    e0 = bethe_ground_energy(n)
    assert abs(r["evals"][0] - e0) <= 1e-10 * abs(e0)
    assert r["rnorms"][0] <= 1e-10
    assert not r["resumed"]
