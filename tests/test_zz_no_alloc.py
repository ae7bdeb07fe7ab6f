"""spined_operator_expectation_device_noalloc (no device allocation; its
dot partials live in a fixed file-scope device array) against the entry it
varies, spined_operator_expectation_device: bit-identical results on two
calls in a row, the second on a smaller input (smaller basis, fewer
columns), real and complex; sizes beyond the array are refused."""
import ctypes as C

import numpy as np
import pytest

import spined
from spined import _lib
from spined.device import DeviceArray
from spined.internal import LatticeSymmetriesException
from helpers import engine, ring

pytestmark = pytest.mark.gpu

CODES = {np.dtype(np.float32): 0, np.dtype(np.float64): 1,
         np.dtype(np.complex64): 2, np.dtype(np.complex128): 3}


def _fn():
    L = _lib.load()
    f = L.spined_operator_expectation_device_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint64,
                  C.c_void_p, C.c_void_p]
    return f


def _both(op, D, cols, dtype, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (cols, D))
    if np.dtype(dtype).kind == "c":
        X = X + 1j * rng.uniform(-1, 1, (cols, D))
    X = np.ascontiguousarray(X.astype(dtype))
    code = CODES[np.dtype(dtype)]
    dx = DeviceArray.from_numpy(X)
    dy = DeviceArray((D,), dtype)
    ref = np.zeros(2 * cols)
    spined.check_status(_lib.load().spined_operator_expectation_device(
        op.ptr, code, D, cols, dx.ptr, D, dy.ptr, ref.ctypes.data))
    got = np.full(2 * cols, np.nan)
    spined.check_status(_fn()(op.ptr.value, code, D, cols, dx.ptr, D, dy.ptr, got.ctypes.data))
    dx.free()
    dy.free()
    return got, ref


@pytest.mark.parametrize("dtype,specs", [
    (np.float64, (ring(20), ring(14))),
    (np.float32, (ring(20), ring(14))),
    (np.complex128, (ring(18, k=5, p=None, z=None), ring(12, k=5, p=None, z=None))),
], ids=["f64", "f32", "c128"])
def test_noalloc_expectation_matches_allocating_entry(dtype, specs):
    cols = (5, 2)
    for i, (spec, c) in enumerate(zip(specs, cols)):
        print("case", spec.name, np.dtype(dtype).name, c, flush=True)
        _, op = engine(spec)
        D = spined.get_number_states(op.basis)
        got, ref = _both(op, D, c, dtype, 100 + i)
        np.testing.assert_array_equal(got, ref)
        assert np.all(np.isfinite(got))


def test_noalloc_expectation_refuses_beyond_its_array():
    _, op = engine(ring(12))
    D = spined.get_number_states(op.basis)
    dy = DeviceArray((D,), np.float64)
    out = np.zeros(2 * 65)
    dx = DeviceArray.zeros((65, D), np.float64)
    st = _fn()(op.ptr.value, 1, D, 65, dx.ptr, D, dy.ptr, out.ctypes.data)
    assert st == 9  # LS_INVALID_ARGUMENT: at most 64 columns
    dx.free()
    dy.free()


def test_noalloc_expectation_refuses_unprepared_operator():
    # a fresh operator on a built basis has not been applied yet: its tables
    # are not on the device, and the call must not upload (allocate) them
    spec = ring(12)
    basis, _ = engine(spec)
    terms = [spined.mk_interaction(m, s) for m, s in spec.terms]
    op2 = spined.mk_operator(basis, terms)
    D = spined.get_number_states(basis)
    dx = DeviceArray.zeros((D,), np.float64)
    dy = DeviceArray((D,), np.float64)
    out = np.zeros(2)
    st = _fn()(op2.ptr.value, 1, D, 1, dx.ptr, D, dy.ptr, out.ctypes.data)
    assert st == 9
    y = spined.apply(op2, np.ones(D))
    assert np.all(np.isfinite(y))
    spined.check_status(_fn()(op2.ptr.value, 1, D, 1, dx.ptr, D, dy.ptr, out.ctypes.data))
    assert out[0] == 0.0 and out[1] == 0.0
    dx.free()
    dy.free()
