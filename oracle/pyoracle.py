"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg load
this module; the engine never does.  The C library is built by
``oracle/Makefile`` (``__graft_entry__.build()`` runs it).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libspined_oracle.so"
_lib = None


def build() -> None:
    import subprocess

    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = C.CDLL(str(_LIB_PATH))
        P = C.c_void_p
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
        L.orc_periodicity.restype = C.c_uint
        L.orc_periodicity.argtypes = [C.c_uint, P]
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P, P, C.POINTER(C.c_int)]
        L.orc_destroy.argtypes = [P]
        for f in ("orc_group_size", "orc_full_group_size", "orc_phase_denominator",
                  "orc_operator_is_real", "orc_number_offdiag_templates"):
            getattr(L, f).restype = C.c_int
            getattr(L, f).argtypes = [P]
        L.orc_get_elements.argtypes = [P, P, P, P]
        L.orc_apply.restype = C.c_uint64
        L.orc_apply.argtypes = [P, C.c_int, C.c_uint64]
        L.orc_orbit.argtypes = [P, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_int),
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_sector_dimension.restype = C.c_double
        L.orc_sector_dimension.argtypes = [P, C.POINTER(C.c_double)]
        L.orc_build.restype = C.c_int64
        L.orc_build.argtypes = [P]
        L.orc_build_range.restype = C.c_int64
        L.orc_build_range.argtypes = [P, C.c_uint64, C.c_uint64]
        L.orc_set_basis.restype = C.c_int
        L.orc_set_basis.argtypes = [P, C.c_int64, u64p]
        L.orc_set_basis_with_stab.restype = C.c_int
        L.orc_set_basis_with_stab.argtypes = [P, C.c_int64, u64p, u32p]
        L.orc_dimension.restype = C.c_int64
        L.orc_dimension.argtypes = [P]
        L.orc_get_basis.argtypes = [P, u64p, u32p]
        L.orc_index.restype = C.c_int64
        L.orc_index.argtypes = [P, C.c_uint64]
        L.orc_add_interaction.restype = C.c_int
        L.orc_add_interaction.argtypes = [P, C.c_int, f64p, C.c_int, P]
        L.orc_matvec_real.restype = C.c_int
        L.orc_matvec_real.argtypes = [P, f64p, f64p, C.c_int64, C.c_int64]
        L.orc_matvec_complex.restype = C.c_int
        L.orc_matvec_complex.argtypes = [P, f64p, f64p, C.c_int64, C.c_int64]
        L.orc_expectation_real.restype = C.c_int
        L.orc_expectation_real.argtypes = [P, f64p, C.POINTER(C.c_double)]
        L.orc_count_offdiag.restype = C.c_int64
        L.orc_count_offdiag.argtypes = [P, C.c_int64, C.c_int64]
        L.orc_random_vector.argtypes = [C.c_uint64, C.c_int64, f64p]
        L.orc_tridiag_eigh.restype = C.c_int
        L.orc_tridiag_eigh.argtypes = [C.c_int, f64p, f64p, f64p, P]
        L.orc_lanczos.restype = C.c_int
        L.orc_lanczos.argtypes = [P, C.c_int, C.c_int, C.c_double, C.c_uint64, P, f64p, P, f64p,
                                  C.POINTER(C.c_int), P, P]
        L.cnn_num_params.restype = C.c_int
        L.cnn_forward.argtypes = [f64p, C.c_int, C.c_int64, u64p, f64p]
        L.cnn_backward.argtypes = [f64p, C.c_int, C.c_int64, u64p, f64p, f64p, f64p]
        _lib = L
    return _lib


def cnn_forward(params, n, states):
    """FP64 (log|psi|, sign logit) per state of the learned-start network."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    s = np.ascontiguousarray(states, dtype=np.uint64)
    out = np.empty(2 * len(s))
    lib().cnn_forward(p, n, len(s), s, out)
    return out[0::2].copy(), out[1::2].copy()


def cnn_backward(params, n, states, ga, gt):
    """FP64 gradient of sum(ga * a + gt * t) with respect to the parameters."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    s = np.ascontiguousarray(states, dtype=np.uint64)
    g = np.zeros(len(p))
    lib().cnn_backward(p, n, len(s), s, np.ascontiguousarray(ga, dtype=np.float64),
                       np.ascontiguousarray(gt, dtype=np.float64), g)
    return g


class OracleError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        super().__init__(f"oracle error {code} {what}")
        self.code = code


def periodicity(perm) -> int:
    p = np.ascontiguousarray(perm, dtype=np.uint32)
    return int(lib().orc_periodicity(len(p), p.ctypes.data))


def random_vector(seed: int, n: int) -> np.ndarray:
    v = np.empty(n, dtype=np.float64)
    lib().orc_random_vector(seed, n, v)
    return v


def tridiag_eigh(alpha, beta, vectors=True):
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    n = len(a)
    if len(b) < max(n - 1, 1):
        b = np.concatenate([b, np.zeros(max(n - 1, 1) - len(b))])
    ev = np.empty(n)
    z = np.empty((n, n)) if vectors else None
    st = lib().orc_tridiag_eigh(n, a, b, ev, z.ctypes.data if vectors else None)
    if st:
        raise OracleError(st, "tridiag")
    # z is stored column-major: row i of the C buffer is eigenvector i
    return (ev, z.T.copy()) if vectors else ev


class System:
    """Group + basis + operator, mirroring the ls_* object chain."""

    def __init__(self, number_spins, hamming_weight=None, spin_inversion=None, symmetries=()):
        L = lib()
        self.n = int(number_spins)
        w = -1 if hamming_weight is None else int(hamming_weight)
        z = 0 if spin_inversion is None else int(spin_inversion)
        gens = list(symmetries)
        perms = np.ascontiguousarray([list(p) for p, _ in gens] or [[0]], dtype=np.uint32)
        secs = np.ascontiguousarray([int(s) for _, s in gens] or [0], dtype=np.uint32)
        if gens and perms.shape[1] != self.n:
            raise OracleError(1, "permutation length")
        st = C.c_int(0)
        self._h = L.orc_create(self.n, w, z, len(gens), perms.ctypes.data, secs.ctypes.data,
                               C.byref(st))
        if not self._h:
            raise OracleError(st.value, "create")
        self._keep = (perms, secs)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    @property
    def group_size(self):
        return lib().orc_group_size(self._h)

    @property
    def full_group_size(self):
        return lib().orc_full_group_size(self._h)

    @property
    def phase_denominator(self):
        return lib().orc_phase_denominator(self._h)

    def elements(self):
        G = self.full_group_size
        perms = np.empty((G, self.n), dtype=np.uint32)
        flips = np.empty(G, dtype=np.int32)
        phases = np.empty(G, dtype=np.int32)
        lib().orc_get_elements(self._h, perms.ctypes.data, flips.ctypes.data, phases.ctypes.data)
        return perms, flips, phases

    def apply(self, g, x):
        return int(lib().orc_apply(self._h, g, x))

    def orbit(self, x):
        r = C.c_uint64()
        p, c, v = C.c_int(), C.c_int(), C.c_int()
        lib().orc_orbit(self._h, x, C.byref(r), C.byref(p), C.byref(c), C.byref(v))
        return r.value, p.value, c.value, bool(v.value)

    def sector_dimension(self):
        im = C.c_double()
        d = lib().orc_sector_dimension(self._h, C.byref(im))
        return d, im.value

    def build(self, first=0, last=(1 << 64) - 1):
        if first == 0 and last == (1 << 64) - 1:
            return int(lib().orc_build(self._h))
        return int(lib().orc_build_range(self._h, first, last))

    def set_basis(self, reps):
        r = np.ascontiguousarray(reps, dtype=np.uint64)
        st = lib().orc_set_basis(self._h, len(r), r)
        if st:
            raise OracleError(st, "set_basis")

    def set_basis_with_stab(self, reps, stab):
        r = np.ascontiguousarray(reps, dtype=np.uint64)
        c = np.ascontiguousarray(stab, dtype=np.uint32)
        st = lib().orc_set_basis_with_stab(self._h, len(r), r, c)
        if st:
            raise OracleError(st, "set_basis_with_stab")

    @property
    def dimension(self):
        return int(lib().orc_dimension(self._h))

    def basis(self):
        D = self.dimension
        reps = np.empty(D, dtype=np.uint64)
        stab = np.empty(D, dtype=np.uint32)
        lib().orc_get_basis(self._h, reps, stab)
        return reps, stab

    def index(self, rep):
        return int(lib().orc_index(self._h, rep))

    def add_interaction(self, matrix, sites):
        m = np.asarray(matrix, dtype=np.complex128)
        n = int(round(np.log2(m.shape[0])))
        if m.shape != (1 << n, 1 << n):
            raise OracleError(9, "matrix shape")
        buf = np.ascontiguousarray(np.stack([m.real, m.imag], axis=-1).reshape(-1))
        s = np.ascontiguousarray(sites, dtype=np.uint32).reshape(-1, n)
        st = lib().orc_add_interaction(self._h, n, buf, s.shape[0], s.ctypes.data)
        if st:
            raise OracleError(st, "interaction")

    @property
    def is_real(self):
        return bool(lib().orc_operator_is_real(self._h))

    def matvec(self, x, rows=None):
        b, e = (0, -1) if rows is None else rows
        D = self.dimension
        n_out = (D if e < 0 else e) - b
        if np.iscomplexobj(x) or not self.is_real:
            xc = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).view(np.float64)
            y = np.empty(2 * n_out)
            st = lib().orc_matvec_complex(self._h, xc, y, b, e)
            if st:
                raise OracleError(st, "matvec")
            return y.view(np.complex128)
        xr = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(n_out)
        st = lib().orc_matvec_real(self._h, xr, y, b, e)
        if st:
            raise OracleError(st, "matvec")
        return y

    def count_offdiag(self, rows=None):
        b, e = (0, -1) if rows is None else rows
        return int(lib().orc_count_offdiag(self._h, b, e))

    def expectation(self, x):
        out = C.c_double()
        st = lib().orc_expectation_real(self._h, np.ascontiguousarray(x, dtype=np.float64),
                                        C.byref(out))
        if st:
            raise OracleError(st, "expectation")
        return out.value

    def lanczos(self, k=1, max_iter=300, tol=1e-10, seed=0, v0=None):
        D = self.dimension
        ev = np.empty(k)
        rn = np.empty(k)
        vecs = np.empty((k, D))
        its = C.c_int()
        m = min(max_iter, D)
        alpha = np.zeros(m)
        beta = np.zeros(m)
        v0a = None if v0 is None else np.ascontiguousarray(v0, dtype=np.float64)
        st = lib().orc_lanczos(self._h, k, max_iter, tol, seed,
                               None if v0a is None else v0a.ctypes.data, ev,
                               vecs.ctypes.data, rn, C.byref(its), alpha.ctypes.data,
                               beta.ctypes.data)
        if st not in (0, 10):
            raise OracleError(st, "lanczos")
        return dict(evals=ev, evecs=vecs.T.copy(), rnorms=rn, iterations=its.value,
                    converged=(st == 0), alpha=alpha[:its.value], beta=beta[:its.value])


def heisenberg_matrix(delta=1.0, j=1.0):
    """Two-site sigma.sigma bond, scaled: [[D,0,0,0],[0,-D,2,0],[0,2,-D,0],[0,0,0,D]]*J."""
    return j * np.array([[delta, 0, 0, 0], [0, -delta, 2, 0], [0, 2, -delta, 0], [0, 0, 0, delta]],
                        dtype=np.float64)
