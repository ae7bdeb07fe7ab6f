"""Objects over the ls_* ABI -- the Python mirror of src/SpinED/Internal.hs.

Names follow the reference (mkSymmetry -> mk_symmetry, getSector ->
get_sector, ...), argument meaning and error behaviour too:
  * host-side validation raises SpinEDException with the reference's
    messages (Internal.hs:109-112, :211-222, :287-290, :324-352, :413-415);
  * a non-zero status from the library raises LatticeSymmetriesException
    (checkStatus, Internal.hs:56-65);
  * every object owns its handle and releases it with the matching
    ls_destroy_* (the ForeignPtr finalisers of Internal.hs:116, :150, :228,
    :362, :402).
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import _lib


class SpinEDException(Exception):
    """Input validation failure (SpinEDException, Internal.hs:34-37)."""


class LatticeSymmetriesException(Exception):
    """Non-zero status from libspined (LatticeSymmetriesException, Internal.hs:29-32)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{message} (code {code})")
        self.code = code
        self.message = message


def check_status(code: int) -> None:
    """checkStatus (Internal.hs:58-65)."""
    if code != 0:
        raise LatticeSymmetriesException(code, _lib.error_message(code))


class _Handle:
    _destroy = None

    def __init__(self, ptr):
        self._ptr = C.c_void_p(ptr)

    @property
    def ptr(self):
        return self._ptr

    def __del__(self):
        p = getattr(self, "_ptr", None)
        if p is not None and p.value:
            getattr(_lib.load(), self._destroy)(p)
            self._ptr = C.c_void_p(None)


def _make(create) -> int:
    """mkObject (Internal.hs:96-97): out-pointer, status check, peek."""
    out = C.c_void_p()
    check_status(create(C.byref(out)))
    return out.value


# ---- symmetries (Internal.hs:67-116) -----------------------------------------
class Symmetry(_Handle):
    _destroy = "ls_destroy_symmetry"

    @property
    def sector(self) -> int:
        return int(_lib.load().ls_get_sector(self.ptr))

    @property
    def periodicity(self) -> int:
        return int(_lib.load().ls_get_periodicity(self.ptr))

    @property
    def phase(self) -> float:
        return float(_lib.load().ls_get_phase(self.ptr))


def mk_symmetry(permutation: Sequence[int], sector: int) -> Symmetry:
    """mkSymmetry (Internal.hs:99-116)."""
    permutation = [int(p) for p in permutation]
    if any(p < 0 for p in permutation):
        raise SpinEDException(f"invalid permutation: {permutation}; indices must be non-negative")
    if sector < 0:
        raise SpinEDException(f"invalid sector: {sector}; expected a non-negative number")
    arr = (C.c_uint * len(permutation))(*permutation)
    L = _lib.load()
    return Symmetry(_make(lambda out: L.ls_create_symmetry(out, len(permutation), arr, int(sector))))


def get_sector(s: Symmetry) -> int:
    return s.sector


def get_periodicity(s: Symmetry) -> int:
    return s.periodicity


def get_phase(s: Symmetry) -> float:
    return s.phase


class SymmetryGroup(_Handle):
    _destroy = "ls_destroy_group"

    @property
    def size(self) -> int:
        return int(_lib.load().ls_get_group_size(self.ptr))


def mk_group(symmetries: Sequence[Symmetry]) -> SymmetryGroup:
    """mkGroup (Internal.hs:141-150)."""
    arr = (C.c_void_p * max(len(symmetries), 1))(*[s.ptr.value for s in symmetries])
    L = _lib.load()
    return SymmetryGroup(_make(lambda out: L.ls_create_group(out, len(symmetries), arr)))


def get_group_size(g: SymmetryGroup) -> int:
    return g.size


# ---- basis (Internal.hs:157-254) ---------------------------------------------
class SpinBasis(_Handle):
    _destroy = "ls_destroy_spin_basis"

    def __init__(self, ptr, group: SymmetryGroup, number_spins: int):
        super().__init__(ptr)
        self.group = group
        self.number_spins = number_spins


def mk_basis(group: SymmetryGroup, number_spins: int, hamming_weight: Optional[int] = None,
             spin_inversion: Optional[int] = None) -> SpinBasis:
    """mkBasis (Internal.hs:199-228)."""
    if number_spins <= 0:
        raise SpinEDException(f"invalid number of spins: {number_spins}; expected a positive number")
    if hamming_weight is None:
        hw = -1
    else:
        if hamming_weight < 0:
            raise SpinEDException(
                f"invalid Hamming weight: {hamming_weight}; expected a non-negative number")
        hw = int(hamming_weight)
    if spin_inversion is None:
        si = 0
    else:
        if spin_inversion not in (1, -1):
            raise SpinEDException(
                f"invalid value for spin inversion: {spin_inversion}; expected either -1 or +1")
        si = int(spin_inversion)
    L = _lib.load()
    ptr = _make(lambda out: L.ls_create_spin_basis(out, group.ptr, int(number_spins), hw, si))
    return SpinBasis(ptr, group, number_spins)


def build_basis(basis: SpinBasis, representatives: Optional[np.ndarray] = None) -> None:
    """buildBasis (Internal.hs:230-236): ls_build, or ls_build_unsafe with reps."""
    L = _lib.load()
    if representatives is not None:
        r = np.ascontiguousarray(representatives, dtype=np.uint64)
        check_status(L.ls_build_unsafe(basis.ptr, len(r), r.ctypes.data))
    else:
        check_status(L.ls_build(basis.ptr))


def get_number_states(basis: SpinBasis) -> int:
    """getNumberStates (Internal.hs:247-254)."""
    out = C.c_uint64()
    check_status(_lib.load().ls_get_number_states(basis.ptr, C.byref(out)))
    return int(out.value)


def basis_get_states(basis: SpinBasis) -> np.ndarray:
    """basisGetStates (Internal.hs:238-244): the representatives as an owned
    array.  One device-to-host copy into the result; ls_get_states, whose
    host buffer the basis keeps for the states object to borrow, serves the
    C ABI (get_states_borrowed)."""
    L = _lib.load()
    reps = C.c_void_p()
    stab = C.c_void_p()
    size = C.c_uint64()
    check_status(L.spined_basis_device_states(basis.ptr, C.byref(reps), C.byref(stab),
                                              C.byref(size)))
    n = int(size.value)
    out = np.empty(n, np.uint64)
    if n:
        check_status(L.spined_memcpy(out.ctypes.data, reps, n * 8, 1))
    return out


def get_states_borrowed(basis: SpinBasis) -> np.ndarray:
    """ls_get_states / ls_states_get_data / ls_destroy_states, as the
    Haskell binding uses them (a copy of the borrowed buffer)."""
    L = _lib.load()
    raw = C.c_void_p()
    check_status(L.ls_get_states(C.byref(raw), basis.ptr))
    try:
        n = int(L.ls_states_get_size(raw))
        data = L.ls_states_get_data(raw)
        out = np.ctypeslib.as_array(data, shape=(n,)).copy() if n else np.zeros(0, np.uint64)
    finally:
        L.ls_destroy_states(raw)
    return out


def basis_get_stabilizers(basis: SpinBasis) -> np.ndarray:
    """|Stab r| per representative (N_r = sqrt(|Stab r| / |G|))."""
    n = get_number_states(basis)
    out = np.empty(n, dtype=np.uint32)
    check_status(_lib.load().spined_basis_get_stabilizers(basis.ptr, out.ctypes.data))
    return out


def sector_dimension(basis: SpinBasis) -> float:
    out = C.c_double()
    check_status(_lib.load().spined_basis_sector_dimension(basis.ptr, C.byref(out)))
    return out.value


# ---- interactions (Internal.hs:256-367) --------------------------------------
class Interaction(_Handle):
    _destroy = "ls_destroy_interaction"

    @property
    def is_real(self) -> bool:
        return bool(_lib.load().ls_interaction_is_real(self.ptr))


def to_matrix(dim: int, rows) -> np.ndarray:
    """toMatrix (Internal.hs:278-291): square check, row-major complex double."""
    rows = [list(r) for r in rows]
    if len(rows) != dim or any(len(r) != dim for r in rows):
        raise SpinEDException(
            f"invalid matrix: {rows}; expected a square matrix of dimension {dim}")
    return np.ascontiguousarray(np.array(rows, dtype=np.complex128))


def mk_interaction(matrix, sites) -> Interaction:
    """mkInteraction' (Internal.hs:324-348) + unsafeMkInteraction (:350-362)."""
    sites = [list(s) for s in sites]
    if not sites:
        raise SpinEDException("zero-point interactions (i.e. constant factors) are not supported")
    n = len(sites[0])
    if n not in (1, 2, 3, 4):
        raise SpinEDException(
            f"currently only 1-, 2-, 3-, and 4-point interactions are supported, but received n={n}")
    if any(len(s) != n for s in sites):
        raise SpinEDException(f"invalid sites: {sites}; expected an array of length-{n} tuples")
    m = to_matrix(1 << n, matrix)
    flat = [int(v) for s in sites for v in s]
    if any(v < 0 for v in flat):
        raise SpinEDException("site indices must be all non-negative numbers")
    arr = np.ascontiguousarray(flat, dtype=np.uint16)
    L = _lib.load()
    create = getattr(L, f"ls_create_interaction{n}")
    mbuf = np.ascontiguousarray(np.stack([m.real, m.imag], axis=-1).reshape(-1))
    return Interaction(_make(lambda out: create(out, mbuf.ctypes.data, len(sites), arr.ctypes.data)))


def is_real_interaction(x: Interaction) -> bool:
    return x.is_real


# ---- operator (Internal.hs:369-458) ------------------------------------------
DTYPE_CODES = {np.dtype(np.float32): 0, np.dtype(np.float64): 1,
               np.dtype(np.complex64): 2, np.dtype(np.complex128): 3}


class Operator_(_Handle):
    """Operator' of the reference: the native operator handle."""

    _destroy = "ls_destroy_operator"

    def __init__(self, ptr, basis: SpinBasis, terms: List[Interaction]):
        super().__init__(ptr)
        self.basis = basis
        self.terms = terms

    @property
    def is_real(self) -> bool:
        return bool(_lib.load().ls_operator_is_real(self.ptr))


def mk_operator(basis: SpinBasis, terms: Sequence[Interaction]) -> Operator_:
    """mkOperator (Internal.hs:391-402)."""
    arr = (C.c_void_p * max(len(terms), 1))(*[t.ptr.value for t in terms])
    L = _lib.load()
    ptr = _make(lambda out: L.ls_create_operator(out, basis.ptr, len(terms), arr))
    return Operator_(ptr, basis, list(terms))


def inplace_apply(op: Operator_, x: np.ndarray, y: np.ndarray) -> None:
    """inplaceApply (Internal.hs:411-429): y <- H x for column-major blocks.

    x and y are (size, block) Fortran-ordered numpy arrays (or 1-D vectors)."""
    if x.shape != y.shape:
        raise SpinEDException(f"dimensions of x and y do not match: {x.shape} != {y.shape}")
    if x.dtype != y.dtype or x.dtype not in DTYPE_CODES:
        raise SpinEDException(f"unsupported datatype {x.dtype}")
    size = x.shape[0]
    block = 1 if x.ndim == 1 else x.shape[1]
    if not (x.flags.f_contiguous and y.flags.f_contiguous):
        raise SpinEDException("x and y must be column-major")
    check_status(_lib.load().ls_operator_matmat(op.ptr, DTYPE_CODES[x.dtype], size, block,
                                                x.ctypes.data, size, y.ctypes.data, size))


def apply(op: Operator_, x: np.ndarray) -> np.ndarray:
    """apply (Internal.hs:431-435)."""
    x = np.asfortranarray(x)
    y = np.empty_like(x, order="F")
    inplace_apply(op, x, y)
    return y


def expectation(op: Operator_, x: np.ndarray) -> np.ndarray:
    """expectation (Internal.hs:437-452): complex <x_j|H|x_j> per column."""
    x = np.asfortranarray(x)
    size = x.shape[0]
    block = 1 if x.ndim == 1 else x.shape[1]
    out = np.zeros(block, dtype=np.complex128)
    check_status(_lib.load().ls_operator_expectation(op.ptr, DTYPE_CODES[x.dtype], size, block,
                                                     x.ctypes.data, size, out.ctypes.data))
    return out


def is_operator_real(op: Operator_) -> bool:
    return op.is_real


# ---- device-resident helpers (spined.h) --------------------------------------
def matvec_device(op: Operator_, x, y, dtype_code: int, y_is_fp64: bool = False,
                  block: int = 1) -> None:
    """y <- H x for device buffers (torch tensors or raw pointers), enqueued on
    the library stream."""
    xp = x if isinstance(x, int) else x.data_ptr()
    yp = y if isinstance(y, int) else y.data_ptr()
    size = get_number_states(op.basis)
    check_status(_lib.load().spined_operator_matmat_device(op.ptr, dtype_code, size, block, xp,
                                                           size, yp, size, int(y_is_fp64)))


def matvec_rows_device(op: Operator_, x, y, dtype_code: int, block: int,
                       y_is_fp64: bool = False) -> None:
    """Y <- H X for a row-interleaved device block (element (row, c) at
    row * block + c), block in {1, 2, 4, 8}; enqueued on the library stream."""
    xp = x if isinstance(x, int) else x.data_ptr()
    yp = y if isinstance(y, int) else y.data_ptr()
    size = get_number_states(op.basis)
    check_status(_lib.load().spined_operator_matmat_rows_device(op.ptr, dtype_code, size, block,
                                                                xp, yp, int(y_is_fp64)))


def synchronize() -> None:
    check_status(_lib.load().spined_synchronize())


def release_device_pool() -> None:
    """Return the pooled device blocks to the driver: the eigensolver's vector
    blocks (otherwise kept for the next solve of the same size) and the entry
    blocks of dropped gather caches (kept for the operator's next build)."""
    check_status(_lib.load().spined_device_release_pool())


def stream_handle() -> int:
    return int(_lib.load().spined_stream() or 0)


def count_offdiag(op: Operator_) -> int:
    out = C.c_uint64()
    check_status(_lib.load().spined_operator_count_offdiag(op.ptr, C.byref(out)))
    return int(out.value)


def set_orbit_method(op: Operator_, method: int) -> None:
    check_status(_lib.load().spined_operator_set_orbit_method(op.ptr, method))


def orbit_method(op: Operator_) -> int:
    return int(_lib.load().spined_operator_orbit_method(op.ptr))


def autotune(op: Operator_, repeats: int = 2) -> dict:
    """Time the matvec over index kind, bucket-table size and row schedule
    (the fastest stays in force) and the block matvec over K = 1, 2, 4, 8."""
    rep = _lib.SpinedAutotuneReport()
    check_status(_lib.load().spined_operator_autotune(op.ptr, repeats, C.byref(rep)))
    out = {f: getattr(rep, f) for f, _ in _lib.SpinedAutotuneReport._fields_}
    out["ms_per_vector_block"] = {k: v for k, v in zip((1, 2, 4, 8), rep.ms_per_vector_block) if v}
    out["ms_per_rows_per_wave"] = {k: v for k, v in zip((64, 32, 16), rep.ms_per_rows_per_wave)
                                   if v}
    return out


def set_rows_per_wave(op: Operator_, rows: int) -> None:
    """Rows per wave of the matrix-free real matvec of exchange-bond
    operators: 64, 32 or 16 (-1 automatic: 64).  Sums are identical."""
    check_status(_lib.load().spined_operator_set_rows_per_wave(op.ptr, rows))


def rows_per_wave(op: Operator_) -> int:
    return _lib.load().spined_operator_get_rows_per_wave(op.ptr)


def set_schedule(op: Operator_, mode: int) -> None:
    """Row schedule of the matvec: -1 automatic, 0 off, 1 on."""
    check_status(_lib.load().spined_operator_set_schedule(op.ptr, mode))


def schedule(op: Operator_) -> int:
    return int(_lib.load().spined_operator_get_schedule(op.ptr))


def set_gather_cache(op: Operator_, mode: int) -> None:
    """Gather cache of a real operator: -1 automatic (built when it fits in
    free device memory), 0 off, 1 on.  Matvec sums are identical either way."""
    check_status(_lib.load().spined_operator_set_gather_cache(op.ptr, mode))


def gather_cache(op: Operator_) -> int:
    """1 while the operator's gather cache is built."""
    return int(_lib.load().spined_operator_get_gather_cache(op.ptr))


def set_index_kind(basis: SpinBasis, kind: int) -> None:
    """0 = bucket table + binary search, 1 = hash table, -1 = automatic."""
    check_status(_lib.load().spined_basis_set_index_kind(basis.ptr, kind))


def index_kind(basis: SpinBasis) -> int:
    return int(_lib.load().spined_basis_get_index_kind(basis.ptr))


def is_complete(basis: SpinBasis) -> int:
    """1: the built basis holds every valid orbit minimum of its sector;
    0: it does not (a partial ls_build_unsafe list); -1: not built."""
    return int(_lib.load().spined_basis_is_complete(basis.ptr))


def set_bucket_bits(basis: SpinBasis, bits: int) -> None:
    check_status(_lib.load().spined_basis_set_bucket_bits(basis.ptr, bits))


def bucket_bits(basis: SpinBasis) -> int:
    return int(_lib.load().spined_basis_get_bucket_bits(basis.ptr))


def debug_orbit(basis: SpinBasis, states, method: int):
    """Host run of the device orbit code (test hook)."""
    s = np.ascontiguousarray(states, dtype=np.uint64)
    n = len(s)
    reps = np.empty(n, np.uint64)
    ph = np.empty(n, np.int32)
    cnt = np.empty(n, np.int32)
    val = np.empty(n, np.uint8)
    check_status(_lib.load().spined_debug_orbit(basis.ptr, method, n, s.ctypes.data,
                                                reps.ctypes.data, ph.ctypes.data,
                                                cnt.ctypes.data, val.ctypes.data))
    return reps, ph, cnt, val.astype(bool)


def debug_orbit_hop(basis: SpinBasis, states, sites):
    """Host run of the nearest-neighbour hop orbit (test hook): the orbit of
    states[q] with its spins at sites[q] and sites[q] + 1 (mod n) exchanged."""
    s = np.ascontiguousarray(states, dtype=np.uint64)
    i = np.ascontiguousarray(sites, dtype=np.int32)
    n = len(s)
    reps = np.empty(n, np.uint64)
    ph = np.empty(n, np.int32)
    cnt = np.empty(n, np.int32)
    val = np.empty(n, np.uint8)
    check_status(_lib.load().spined_debug_orbit_hop(basis.ptr, n, s.ctypes.data, i.ctypes.data,
                                                    reps.ctypes.data, ph.ctypes.data,
                                                    cnt.ctypes.data, val.ctypes.data))
    return reps, ph, cnt, val.astype(bool)


def ring_available(basis: SpinBasis) -> bool:
    return bool(_lib.load().spined_debug_ring_available(basis.ptr))


def torus_shape(basis: SpinBasis):
    """(l1, l2) of the generic orbit's torus factorisation; (0, 0) when off."""
    l1, l2 = C.c_int32(), C.c_int32()
    check_status(_lib.load().spined_debug_torus_shape(basis.ptr, C.byref(l1), C.byref(l2)))
    return l1.value, l2.value
