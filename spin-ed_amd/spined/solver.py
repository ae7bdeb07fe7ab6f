"""spined_eigsh wrapper: the device-resident replacement of PRIMME's eigh
(src/SpinED.hs:383-404)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import _lib
from .internal import Operator_, check_status, get_number_states

PRECISION = {"fp64": 0, "mixed": 1}
START = {"random": 0, "vector": 1, "learned": 2}


@dataclass
class EigshResult:
    evals: np.ndarray
    evecs: Optional[np.ndarray]   # (D, k) column-major, or None
    rnorms: np.ndarray
    status: int
    stats: dict


def eigsh(op: Operator_, k: int = 1, tol: float = 1e-10, *, precision: str = "fp64",
          start: str = "random", seed: int = 0x5EED, start_vector: Optional[np.ndarray] = None,
          max_basis_size: int = 0, min_restart_size: int = 0, max_block_size: int = 0,
          max_matvecs: int = 0, want_vectors: bool = True, dtype=np.float64,
          verbose: bool = False, monitor: Optional[Callable[[dict], bool]] = None,
          model=None, sign_mask: int = 0,
          checkpoint: Optional[Callable[[np.ndarray, dict], None]] = None,
          checkpoint_every: int = 0) -> EigshResult:
    """... checkpoint(v, info) receives, after every checkpoint_every-th
    restart, the lowest Ritz vector of the current basis as a host array
    (float32 in the FP32-storage phase, else float64, complex128 for complex
    sectors) and the monitor's info; start_vector=v resumes from it."""
    L = _lib.load()
    D = get_number_states(op.basis)
    opts = _lib.SpinedEigshOptions()
    L.spined_eigsh_options_init(C.byref(opts))
    opts.precision = PRECISION[precision]
    opts.start = START[start]
    opts.seed = seed
    opts.max_basis_size = max_basis_size
    opts.min_restart_size = min_restart_size
    opts.max_block_size = max_block_size
    opts.max_matvecs = max_matvecs
    opts.verbose = int(verbose)
    sv = None
    if model is not None:
        opts.start = START["learned"]
        opts.model = model.handle.value
        opts.start_sign_mask = sign_mask
    if start_vector is not None:
        sv = np.ascontiguousarray(start_vector, dtype=np.float64)
        if sv.shape != (D,):
            raise ValueError("start_vector must have length = dimension")
        opts.start_vector = sv.ctypes.data
        opts.start = START["vector"]
    evals = np.zeros(k)
    rn = np.zeros(k)
    dtype = np.dtype(dtype)
    if not op.is_real and dtype.kind != "c":
        # complex (momentum) sectors: Hermitian Lanczos, complex eigenvectors
        dtype = np.dtype(np.complex128 if dtype == np.float64 else np.complex64)
    code = {np.dtype(np.float64): 1, np.dtype(np.float32): 0,
            np.dtype(np.complex64): 2, np.dtype(np.complex128): 3}[dtype]
    evecs = np.zeros((D, k), dtype=dtype, order="F") if want_vectors else None

    def _mon(info_p, _ctx):
        if monitor is None:
            return 0
        info = info_p.contents
        d = dict(matvecs=info.matvecs, restarts=info.restarts, phase=info.phase,
                 nconverged=info.nconverged,
                 ritz_values=[info.ritz_values[i] for i in range(info.k)],
                 residuals=[info.residuals[i] for i in range(info.k)])
        return 1 if monitor(d) else 0

    cb = _lib.MONITOR_FN(_mon)

    def _cp(ptr, code, n, info_p, _ctx):
        dt = {0: np.float32, 1: np.float64, 3: np.complex128}[int(code)]
        v = np.empty(int(n), dtype=dt)
        check_status(L.spined_memcpy(v.ctypes.data, ptr, v.nbytes, 1))
        info = info_p.contents
        checkpoint(v, dict(matvecs=info.matvecs, restarts=info.restarts, phase=info.phase,
                           nconverged=info.nconverged,
                           ritz_values=[info.ritz_values[i] for i in range(info.k)],
                           residuals=[info.residuals[i] for i in range(info.k)]))

    cpf = _lib.CHECKPOINT_FN(_cp) if checkpoint is not None else _lib.CHECKPOINT_FN()
    if checkpoint is not None and checkpoint_every > 0:
        opts.checkpoint = cpf
        opts.checkpoint_every = checkpoint_every
    status = L.spined_eigsh(op.ptr, code, k, tol, C.byref(opts), evals.ctypes.data,
                            evecs.ctypes.data if evecs is not None else None, D,
                            rn.ctypes.data, cb, None)
    if status not in (0, 10):
        check_status(status)
    st = _lib.SpinedStats()
    check_status(L.spined_get_last_stats(op.ptr, C.byref(st)))
    stats = {f: getattr(st, f) for f, _ in _lib.SpinedStats._fields_}
    return EigshResult(evals, evecs, rn, status, stats)
