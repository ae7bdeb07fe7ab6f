"""Complex (momentum) sectors end to end: spined_eigsh with the complex
Hermitian Lanczos against dense LAPACK diagonalisation of the oracle's
sector matrix (built column by column from the oracle's complex matvec),
residuals through the oracle's matvec, and the k <-> N - k symmetry of the
ring spectrum (complex-conjugate sectors have identical spectra)."""
import numpy as np
import pytest

import spined
from helpers import engine, oracle, ring

pytestmark = pytest.mark.gpu


def _dense(o):
    D = o.dimension
    H = np.empty((D, D), dtype=np.complex128)
    for j in range(D):
        e = np.zeros(D, dtype=np.complex128)
        e[j] = 1.0
        H[:, j] = o.matvec(e)
    return H


@pytest.mark.parametrize("n,k,z,nev", [(12, 5, None, 3), (14, 3, 1, 2), (16, 3, None, 4)])
def test_complex_sector_vs_dense(n, k, z, nev):
    spec = ring(n, k=k, p=None, z=z, name=f"ring{n}_k{k}_z{z}")
    print("case", spec.name, flush=True)
    basis, op = engine(spec)
    assert not op.is_real
    o = oracle(spec)
    H = _dense(o)
    assert np.max(np.abs(H - H.conj().T)) <= 1e-12
    ref = np.linalg.eigvalsh(H)[:nev]
    r = spined.eigsh(op, k=nev, tol=1e-10)
    print(r.evals, ref, r.rnorms, r.stats, flush=True)
    assert r.status == 0
    assert r.evecs.dtype == np.complex128
    np.testing.assert_allclose(r.evals, ref, rtol=1e-10, atol=1e-10)
    assert np.all(r.rnorms <= 1e-10)
    for i in range(nev):
        v = r.evecs[:, i]
        assert abs(np.linalg.norm(v) - 1) <= 1e-10
        res = np.linalg.norm(o.matvec(v) - r.evals[i] * v)
        assert res <= 1e-9, (i, res)


def test_complex64_eigenvectors():
    spec = ring(12, k=5, p=None, z=None, name="ring12_k5")
    basis, op = engine(spec)
    r = spined.eigsh(op, k=1, tol=1e-10, dtype=np.complex64)
    assert r.evecs.dtype == np.complex64
    o = oracle(spec)
    v = r.evecs[:, 0].astype(np.complex128)
    assert np.linalg.norm(o.matvec(v) - r.evals[0] * v) <= 1e-5


def test_conjugate_sectors_agree_ring24():
    """Sectors k and N - k of the N = 24 ring: identical spectra."""
    e = []
    for k in (5, 19):
        spec = ring(24, k=k, p=None, z=None, name=f"ring24_k{k}")
        basis, op = engine(spec)
        r = spined.eigsh(op, k=2, tol=1e-10)
        assert r.status == 0 and np.all(r.rnorms <= 1e-10)
        e.append(r.evals)
        o = oracle(spec)
        v = r.evecs[:, 0]
        assert np.linalg.norm(o.matvec(v) - r.evals[0] * v) <= 1e-9
    np.testing.assert_allclose(e[0], e[1], rtol=1e-10, atol=0)


def test_learned_start_rejected_for_complex_sector():
    spec = ring(12, k=5, p=None, z=None, name="ring12_k5")
    basis, op = engine(spec)
    from spined import learned
    model = learned.CNN(learned.init_params(1))
    with pytest.raises(Exception):
        spined.eigsh(op, k=1, model=model)
