"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars: representatives and stabilisers bit-exact; FP64 matvec within 4 ulp
per row (same summation order and fma pattern as the oracle, so mostly
bit-exact); FP32-storage matvec within FP32 rounding of x; eigenvalues within
1e-10 relative of the oracle / closed forms; residual <= tol; eigenvector
overlap >= 1 - 1e-8.
"""
import numpy as np
import pytest

import pyoracle
import spined
from spined import lattice
from spined.internal import LatticeSymmetriesException
from helpers import Spec, engine, oracle, ring, torus
from test_oracle_pins import bethe_ground_energy

pytestmark = pytest.mark.gpu

BASIS_CASES = [
    Spec(10, None, None, [], [(lattice.HEISENBERG, lattice.ring_bonds(10))], "chain10_full"),
    ring(10, k=5, p=1, z=-1, name="chain10_example"),          # example/heisenberg_chain_10.yaml
    ring(16, k=0, p=0, z=1),
    ring(24, k=0, p=0, z=1),
    ring(20, k=3, p=None, z=None, name="ring20_k3_complex"),
    ring(18, k=9, p=1, z=-1),
    ring(14, k=None, p=None, z=1, name="xxz14_flip"),
    Spec(16, 8, 1, lattice.square_symmetries(4), [(lattice.HEISENBERG, lattice.square_bonds(4))], "square4x4"),
    Spec(12, None, None, [(lattice.ring_translation(12), 2)], [(lattice.HEISENBERG, lattice.ring_bonds(12))], "ring12_now_k2"),
    # the generic orbit's torus factorisation: complex momenta, reflections, flips
    torus(4, 3, kx=1, ky=2),
    torus(3, 4, reflect=True, z=-1),
    torus(4, 4, kx=2, ky=0, z=1),
    torus(6, 3, kx=0, ky=1, reflect=True),
    Spec(25, 12, None, lattice.square_symmetries(5), [(lattice.HEISENBERG, lattice.square_bonds(5))],
         "square5x5"),
]


@pytest.mark.parametrize("spec", BASIS_CASES, ids=lambda s: s.name)
def test_basis_bit_exact(spec):
    print("case", spec.name, flush=True)
    basis, _ = engine(spec)
    o = oracle(spec)
    reps = spined.basis_get_states(basis)
    oreps, ostab = o.basis()
    assert reps.dtype == np.uint64
    np.testing.assert_array_equal(reps, oreps)
    # the borrowed buffer of ls_get_states (the Haskell binding's path)
    np.testing.assert_array_equal(spined.get_states_borrowed(basis), oreps)
    assert spined.get_number_states(basis) == len(oreps)
    d, _ = o.sector_dimension()
    assert len(oreps) == round(d)
    if spec.name.startswith(("torus", "square")):
        assert spined.torus_shape(basis) != (0, 0)


def _ulp_diff(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    sp = np.spacing(np.maximum(np.abs(a), np.abs(b)))
    return np.max(np.abs(a - b) / np.where(sp > 0, sp, 1.0)) if len(a) else 0.0


@pytest.mark.parametrize("spec", BASIS_CASES, ids=lambda s: s.name)
def test_matvec_fp64(spec):
    print("case", spec.name, flush=True)
    basis, op = engine(spec)
    o = oracle(spec)
    D = o.dimension
    rng = np.random.default_rng(3)
    for method in ([0, 1] if spined.ring_available(basis) else [0]):
        spined.set_orbit_method(op, method)
        if o.is_real:
            x = rng.uniform(-1, 1, D)
            y = spined.apply(op, x)
            yr = o.matvec(x)
            scale = np.max(np.abs(yr)) if D else 1.0
            assert np.max(np.abs(y - yr), initial=0) <= 1e-13 * scale, (method,)
            print("method", method, "max ulp", _ulp_diff(y, yr))
        else:
            x = rng.uniform(-1, 1, D) + 1j * rng.uniform(-1, 1, D)
            y = spined.apply(op, x)
            yr = o.matvec(x)
            scale = np.max(np.abs(yr)) if D else 1.0
            assert np.max(np.abs(y - yr), initial=0) <= 1e-13 * scale, (method,)


def test_matvec_fp32_and_mixed():
    from spined.device import DeviceArray
    spec = ring(24)
    basis, op = engine(spec)
    o = oracle(spec)
    D = o.dimension
    x = np.random.default_rng(5).uniform(-1, 1, D).astype(np.float32)
    yr = o.matvec(x.astype(np.float64))   # exact-in-x reference
    y32 = spined.apply(op, x)
    assert y32.dtype == np.float32
    assert np.max(np.abs(y32 - yr)) <= 1e-6 * np.max(np.abs(yr))
    xd = DeviceArray.from_numpy(x)
    yd = DeviceArray.zeros(D, np.float64)
    spined.matvec_device(op, xd, yd, 0, y_is_fp64=True)
    spined.synchronize()
    ym = yd.numpy()
    # FP32 storage, FP64 accumulation: identical to the FP64 oracle on the same x
    assert np.max(np.abs(ym - yr)) <= 1e-13 * np.max(np.abs(yr))


def test_eigsh_config1_chain10_full():
    spec = BASIS_CASES[0]
    basis, op = engine(spec)
    r = spined.eigsh(op, k=1, tol=1e-10)
    e_exact = -18.061785417968172  # dense ED and Bethe ansatz
    assert r.status == 0
    assert abs(r.evals[0] - e_exact) <= 1e-10 * abs(e_exact)
    assert r.rnorms[0] <= 1e-10
    o = oracle(spec)
    ref = o.lanczos(k=1, tol=1e-11)
    ov = abs(np.dot(r.evecs[:, 0], ref["evecs"][:, 0]))
    assert ov >= 1 - 1e-8


def test_eigsh_config2_ring24_lowest4():
    spec = ring(24)
    basis, op = engine(spec)
    r = spined.eigsh(op, k=4, tol=1e-10)
    o = oracle(spec)
    ref = o.lanczos(k=4, tol=1e-11, max_iter=400)
    print("gpu", r.evals, r.rnorms, r.stats)
    print("cpu", ref["evals"], ref["rnorms"])
    assert r.status == 0
    np.testing.assert_allclose(r.evals, ref["evals"], rtol=1e-10, atol=0)
    assert abs(r.evals[0] - (-42.680058066148944)) <= 1e-10 * 42.68  # Bethe ansatz
    assert np.all(r.rnorms <= 1e-10)
    for i in range(4):
        ov = abs(np.dot(r.evecs[:, i], ref["evecs"][:, i]))
        assert ov >= 1 - 1e-8, (i, ov)


def test_eigsh_mixed_precision_ring24():
    spec = ring(24)
    basis, op = engine(spec)
    r64 = spined.eigsh(op, k=1, tol=1e-10)
    rm = spined.eigsh(op, k=1, tol=1e-10, precision="mixed")
    print("fp64", r64.evals, r64.stats)
    print("mixed", rm.evals, rm.stats)
    assert rm.status == 0
    assert abs(rm.evals[0] - r64.evals[0]) <= 1e-10 * abs(r64.evals[0])
    assert rm.rnorms[0] <= 1e-10
    assert rm.stats["matvecs_mixed"] > 0
    assert abs(np.dot(rm.evecs[:, 0], r64.evecs[:, 0])) >= 1 - 1e-8


def test_operator_before_build_binds_lazily():
    # SpinED creates the operator before ls_build (SpinED.hs:243-248, :315-337)
    spec = ring(16)
    basis, op = engine(spec, build=False)
    spined.build_basis(basis)
    o = oracle(spec)
    x = np.random.default_rng(1).uniform(-1, 1, o.dimension)
    np.testing.assert_allclose(spined.apply(op, x), o.matvec(x), rtol=0, atol=1e-12)


def test_build_unsafe_roundtrip():
    spec = ring(16)
    b1, _ = engine(spec)
    reps = spined.basis_get_states(b1)
    b2, op2 = engine(spec, build=False)
    spined.build_basis(b2, reps)
    np.testing.assert_array_equal(spined.basis_get_states(b2), reps)
    o = oracle(spec)
    x = np.random.default_rng(2).uniform(-1, 1, o.dimension)
    np.testing.assert_allclose(spined.apply(op2, x), o.matvec(x), rtol=0, atol=1e-12)


def test_build_unsafe_refuses_a_built_basis():
    """A built basis keeps its list: operators bound to it hold row schedules
    and gather caches sized for it, so ls_build_unsafe on it is refused (and
    ls_build stays a no-op), and the operator still matches the oracle."""
    spec = ring(16)
    basis, op = engine(spec)
    o = oracle(spec)
    x = np.random.default_rng(3).uniform(-1, 1, o.dimension)
    y0 = spined.apply(op, x)
    reps = spined.basis_get_states(basis)
    with pytest.raises(LatticeSymmetriesException) as e:
        spined.build_basis(basis, reps[: len(reps) // 2])
    assert e.value.code == 9
    spined.build_basis(basis)
    np.testing.assert_array_equal(spined.basis_get_states(basis), reps)
    np.testing.assert_array_equal(spined.apply(op, x), y0)
    np.testing.assert_allclose(y0, o.matvec(x), rtol=0, atol=1e-12)


@pytest.mark.parametrize("spec,dtype", [
    (ring(16), np.float64), (ring(16), np.float32),
    (ring(14, k=3, p=None, z=None, name="ring14_k3"), np.complex128),
    (ring(14, k=3, p=None, z=None, name="ring14_k3"), np.complex64),
], ids=["f64", "f32", "c128", "c64"])
def test_expectation_on_device_matches_oracle(spec, dtype):
    """ls_operator_expectation: <x_c|H|x_c> per column, H x_c and the FP64 dot
    on the device (Internal.hs:437-452), against numpy's dot of the oracle's
    H x_c; also the device-pointer entry point on the same columns."""
    from spined import _lib
    from spined.device import DeviceArray
    _, op = engine(spec)
    o = oracle(spec)
    D = o.dimension
    rng = np.random.default_rng(11)
    cplx = np.dtype(dtype).kind == "c"
    X = rng.uniform(-1, 1, (D, 3))
    if cplx:
        X = X + 1j * rng.uniform(-1, 1, (D, 3))
    X = np.asfortranarray(X.astype(dtype))
    got = spined.expectation(op, X)
    tol = 1e-12 if np.dtype(dtype) in (np.float64, np.complex128) else 1e-5
    ref = []
    for c in range(3):
        xc = X[:, c].astype(np.complex128 if cplx else np.float64)
        ref.append(np.vdot(xc, o.matvec(xc)))
    ref = np.array(ref)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= tol * scale, (got, ref)
    code = {np.dtype(np.float32): 0, np.dtype(np.float64): 1,
            np.dtype(np.complex64): 2, np.dtype(np.complex128): 3}[np.dtype(dtype)]
    dx = DeviceArray.from_numpy(np.ascontiguousarray(X.T))
    dy = DeviceArray((D,), dtype)
    out = np.zeros(6)
    spined.check_status(_lib.load().spined_operator_expectation_device(
        op.ptr, code, D, 3, dx.ptr, D, dy.ptr, out.ctypes.data))
    np.testing.assert_array_equal(out[0::2] + 1j * out[1::2], np.asarray(got, dtype=np.complex128))


@pytest.mark.parametrize("want_vectors", [True, False])
def test_eigsh_mixed_ring32(want_vectors):
    """FP32 storage on a basis of more than 2^20 states (the FP32 phase runs
    to its floor, then the FP64 phase finishes): E0 equals the FP64 solve and
    the Bethe-ansatz value, the residual is explicit, and the eigenvector
    overlaps the FP64 one to 1 - 1e-8."""
    spec = ring(32)
    basis, op = engine(spec)
    rm = spined.eigsh(op, k=1, tol=1e-10, precision="mixed", want_vectors=want_vectors,
                      verbose=True)
    r64 = spined.eigsh(op, k=1, tol=1e-10, want_vectors=want_vectors)
    print("fp64", r64.evals, r64.rnorms, r64.stats)
    print("mixed", rm.evals, rm.rnorms, rm.stats)
    assert rm.status == 0 and r64.status == 0
    assert rm.rnorms[0] <= 1e-10
    e_bethe = bethe_ground_energy(32)
    assert abs(rm.evals[0] - r64.evals[0]) <= 1e-10 * abs(r64.evals[0])
    assert abs(rm.evals[0] - e_bethe) <= 1e-10 * abs(e_bethe)
    assert rm.stats["matvecs_mixed"] > 0
    if want_vectors:
        assert abs(np.linalg.norm(rm.evecs[:, 0]) - 1) <= 1e-12
        assert abs(np.dot(rm.evecs[:, 0], r64.evecs[:, 0])) >= 1 - 1e-8
