"""spined -- MI355X-native exact diagonalisation behind SpinED's ls_* ABI.

The compute path is libspined.so (hand-written gfx950 HIP kernels behind the
ls_* C ABI of include/lattice_symmetries.h and the spined_* entries of
include/spined.h).  This package is the host side: ctypes bindings that
mirror src/SpinED/Internal.hs (internal.py), the eigensolver wrapper
(solver.py) and the SpinED application flow (config/app/hdf5)."""
from .internal import *  # noqa: F401,F403
from .internal import SpinEDException, LatticeSymmetriesException  # noqa: F401
from .solver import eigsh, EigshResult  # noqa: F401
from . import lattice  # noqa: F401
