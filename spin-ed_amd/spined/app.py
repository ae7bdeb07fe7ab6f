"""The SpinED application flow -- mirror of src/SpinED.hs (orchestration) and
app/Main.hs (runApp), on top of libspined.

run_app(input.yaml) performs, in the reference's order (app/Main.hs:14-27):
  readConfig -> toConfig -> prepareOutputFile -> buildRepresentatives ->
  isOperatorReal -> withDatatype -> diagonalize -> computeExpectations.
The one difference is inside `diagonalize`: the eigensolve runs on the GPU
through spined_eigsh instead of PRIMME's eigh with a host callback
(SpinED.hs:404).
"""
from __future__ import annotations

import argparse
import logging
import sys
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import hdf5, internal as I
from .config import ConfigSpec, OperatorSpec, read_config
from .solver import eigsh

log = logging.getLogger("spined")

BASIS_PATH = "/basis"
HAMILTONIAN_PATH = "/hamiltonian"
OBSERVABLES_PATH = "/observables"
WORKSPACE_PATH = "/_workspace"
# Lanczos checkpoint (no counterpart in the reference, which creates
# /_workspace and leaves it empty): the lowest Ritz vector of the current
# basis, its Ritz values and the matvecs so far, rewritten every
# checkpoint_every restarts; a later run on the same output file starts
# from it (ground-state runs, real sectors)
CHECKPOINT_PATH = WORKSPACE_PATH + "/lanczos"
DEFAULT_CHECKPOINT_EVERY = 8


@dataclass
class Operator:
    """Operator (SpinED.hs:193-196): a name and the native operator."""
    name: str
    obj: I.Operator_


@dataclass
class UserConfig:
    """UserConfig (SpinED.hs:202-213)."""
    basis: I.SpinBasis
    hamiltonian: Operator
    observables: List[Operator]
    output: str
    num_evals: int
    eps: float
    datatype: str
    max_basis_size: int
    max_block_size: int
    min_restart_size: int
    spec: ConfigSpec


def to_symmetry(s):
    return I.mk_symmetry(s.permutation, s.sector)


def to_basis(spec) -> I.SpinBasis:
    """toBasis (SpinED.hs:178-188)."""
    if spec.number_spins > 64:
        raise I.SpinEDException(
            f"invalid number_spins: {spec.number_spins}; exact diagonalization is not feasible "
            "for systems larger than 64 spins")
    log.debug("Building symmetry group...")
    group = I.mk_group([to_symmetry(s) for s in spec.symmetries])
    log.info("Symmetry group contains %d elements", group.size)
    return I.mk_basis(group, spec.number_spins, spec.hamming_weight, spec.spin_inversion)


def to_interaction(spec) -> I.Interaction:
    return I.mk_interaction(spec.matrix, spec.sites)


def to_operator(basis: I.SpinBasis, spec: OperatorSpec) -> Operator:
    """toOperator (SpinED.hs:198-200)."""
    return Operator(spec.name, I.mk_operator(basis, [to_interaction(t) for t in spec.terms]))


def to_config(spec: ConfigSpec) -> UserConfig:
    """toConfig (SpinED.hs:243-248)."""
    basis = to_basis(spec.basis)
    ham = to_operator(basis, spec.hamiltonian)
    obs = [to_operator(basis, o) for o in spec.observables]
    return UserConfig(basis, ham, obs, spec.output, spec.number_vectors, spec.precision,
                      spec.datatype, spec.max_primme_basis_size, spec.max_primme_block_size,
                      spec.min_primme_restart_size, spec)


def prepare_output_file(cfg: UserConfig) -> None:
    """prepareOutputFile (SpinED.hs:289-294)."""
    with hdf5.File(cfg.output, "a") as f:
        for g in (BASIS_PATH, HAMILTONIAN_PATH, OBSERVABLES_PATH, WORKSPACE_PATH):
            f.create_group(g)


class _PendingWrite:
    """A dataset written by a background thread while the GPU works (libhdf5
    runs outside the GIL); wait() joins it and re-raises its error.  Only
    one thread touches the file at a time: every other access of the output
    file waits for it first."""

    def __init__(self, path, dataset, data):
        import threading
        self._err = None

        def run():
            try:
                with hdf5.File(path, "a") as f:
                    f.write(dataset, data)
            except BaseException as e:  # re-raised by wait()
                self._err = e
        self._t = threading.Thread(target=run, daemon=True)
        self._t.start()

    def wait(self) -> None:
        if self._t is not None:
            self._t.join()
            self._t = None
            if self._err is not None:
                raise self._err


def build_representatives(cfg: UserConfig, background: bool = False):
    """buildRepresentatives (SpinED.hs:315-337): reload /basis/representatives
    when present (ls_build_unsafe), else build and store them.  background:
    return the representatives still to be stored (None when loaded) and
    leave the store to the caller."""
    reps = None
    with hdf5.File(cfg.output, "a") as f:
        if f.exists(BASIS_PATH + "/representatives"):
            log.info("Loading representatives from %s/representatives...", BASIS_PATH)
            reps = f.read(BASIS_PATH + "/representatives", np.uint64)
    if reps is None:
        log.info("Building a list of representatives...")
    I.build_basis(cfg.basis, reps)
    dim = I.get_number_states(cfg.basis)
    log.info("Hilbert space dimension is %d", dim)
    if reps is not None:
        return None
    states = I.basis_get_states(cfg.basis)
    if background:
        return states
    with hdf5.File(cfg.output, "a") as f:
        f.write(BASIS_PATH + "/representatives", states)
    return None


def with_datatype(is_real: bool, datatype: str):
    """withDatatype (SpinED.hs:344-352)."""
    if is_real:
        return np.float32 if datatype == "float32" else np.float64
    return np.complex64 if datatype == "float32" else np.complex128


def read_checkpoint(cfg: UserConfig, dim: int):
    """The start vector of an earlier run's checkpoint, or None: the
    converged eigenvector of a finished run (marked in /_workspace/lanczos,
    stored once, in /hamiltonian/eigenvectors), else the Ritz vector of an
    interrupted one."""
    with hdf5.File(cfg.output, "a") as f:
        if (f.exists(CHECKPOINT_PATH + "/converged")
                and f.exists(HAMILTONIAN_PATH + "/eigenvectors")):
            v = hdf5.read_eigenvectors(f, dim, np.float64)
            v = v[:, 0] if v.ndim == 2 and v.shape[0] == dim else np.zeros(0)
        elif f.exists(CHECKPOINT_PATH + "/start_vector"):
            v = f.read(CHECKPOINT_PATH + "/start_vector", np.float64)
        else:
            return None
    return v if v.shape == (dim,) and np.all(np.isfinite(v)) and np.any(v) else None


_UNREAD = object()


def diagonalize(cfg: UserConfig, dtype, model=None,
                checkpoint_every: int = DEFAULT_CHECKPOINT_EVERY, wait_io=lambda: None,
                resume=_UNREAD):
    """diagonalize (SpinED.hs:370-411): lowest number_vectors eigenpairs,
    written to /hamiltonian/{eigenvalues, eigenvectors, residuals}; with
    checkpoint_every > 0 the solver's lowest Ritz vector goes to
    /_workspace/lanczos every checkpoint_every restarts, and a checkpoint
    found there (same dimension, one eigenpair, real sector) is the start
    (resume: that vector or None, when the caller has read it already).
    wait_io() is called before every access of the output file."""
    dtype = np.dtype(dtype)
    log.debug("Using max_primme_basis_size=%d, max_primme_block_size=%d, "
              "min_primme_restart_size=%d", cfg.max_basis_size, cfg.max_block_size,
              cfg.min_restart_size)
    log.info("Diagonalizing %s...", cfg.hamiltonian.name)

    def monitor(info):
        log.info("restart %d matvecs %d ritz %s residuals %s", info["restarts"], info["matvecs"],
                 info["ritz_values"], info["residuals"])
        return False

    real = np.dtype(dtype).kind != "c"
    dim = I.get_number_states(cfg.basis)
    start_vector = None
    net, sign_mask = None, 0
    resumed = None
    if real and cfg.num_evals == 1:
        if resume is _UNREAD:
            wait_io()
            resumed = read_checkpoint(cfg, dim)
        else:
            resumed = resume
    if resumed is not None:
        log.info("Resuming from the checkpoint in %s...", CHECKPOINT_PATH)
        start_vector = resumed
    elif cfg.spec.start_vector == "learned":
        # evaluated on the device by the solver: the vector never leaves it
        from . import learned
        net = learned.model_for(cfg, model)
        sign_mask = learned.marshall_mask(cfg.spec.basis.number_spins)

    def write_checkpoint(v, info):
        wait_io()
        with hdf5.File(cfg.output, "a") as f:
            f.delete(CHECKPOINT_PATH + "/converged")
            f.write(CHECKPOINT_PATH + "/start_vector", v.astype(np.float64), warn=False)
            f.write(CHECKPOINT_PATH + "/ritz_values", np.asarray(info["ritz_values"]), warn=False)
            f.write(CHECKPOINT_PATH + "/matvecs", np.array([info["matvecs"]], np.int64),
                    warn=False)

    res = eigsh(cfg.hamiltonian.obj, k=cfg.num_evals, tol=cfg.eps if cfg.eps > 0 else 1e-10,
                precision=cfg.spec.solver_precision, seed=cfg.spec.seed,
                start_vector=start_vector, model=net, sign_mask=sign_mask,
                max_basis_size=cfg.max_basis_size,
                min_restart_size=cfg.min_restart_size, max_block_size=cfg.max_block_size,
                dtype=dtype, monitor=monitor,
                checkpoint=write_checkpoint if real and checkpoint_every > 0 else None,
                checkpoint_every=checkpoint_every)
    if res.status != 0:
        log.warning("eigensolver stopped before reaching the tolerance: residuals %s", res.rnorms)
    log.info("Obtained eigenvalues %s", res.evals)
    real_t = np.float32 if dtype in (np.float32, np.complex64) else np.float64
    evals = res.evals.astype(real_t)
    rnorms = res.rnorms.astype(real_t)
    wait_io()
    with hdf5.File(cfg.output, "a") as f:
        f.write(HAMILTONIAN_PATH + "/eigenvalues", evals)
        # KnownDataset (Block a): the column-major D x k block as a row-major (k, D) matrix
        f.write(HAMILTONIAN_PATH + "/eigenvectors", np.ascontiguousarray(res.evecs.T))
        f.write(HAMILTONIAN_PATH + "/residuals", rnorms)
        if real and checkpoint_every > 0 and res.status == 0:
            # converged: a later run on this file starts from the eigenvector
            # just written (the interrupted-run vector is no longer needed)
            f.delete(CHECKPOINT_PATH + "/start_vector")
            f.write(CHECKPOINT_PATH + "/converged", np.array([1], np.int64), warn=False)
            f.write(CHECKPOINT_PATH + "/ritz_values", np.asarray(res.evals), warn=False)
            f.write(CHECKPOINT_PATH + "/matvecs", np.array([res.stats["matvecs"]], np.int64),
                    warn=False)
    diagonalize.last = dict(matvecs=int(res.stats["matvecs"]), resumed=resumed is not None)
    return evals, res.evecs, rnorms


def compute_expectations(cfg: UserConfig, evecs: np.ndarray) -> None:
    """computeExpectations / computeExpectation (SpinED.hs:296-313)."""
    for o in cfg.observables:
        log.info("Computing expectation values of %s...", o.name)
        x = evecs
        if not I.is_operator_real(o.obj) and not np.iscomplexobj(evecs):
            # a complex observable in a real sector: promote the eigenvectors
            x = evecs.astype(np.complex64 if evecs.dtype == np.float32 else np.complex128)
        m = I.expectation(o.obj, np.asfortranarray(x))
        log.info("Obtained expectation values %s", m)
        with hdf5.File(cfg.output, "a") as f:
            f.write(OBSERVABLES_PATH + "/" + o.name, m.astype(np.complex128))


def run_app(path: str, debug: bool = False,
            checkpoint_every: int = DEFAULT_CHECKPOINT_EVERY) -> dict:
    """runApp (app/Main.hs:14-27)."""
    if debug:
        I._lib.load().ls_enable_logging()
    t0 = time.perf_counter()
    cfg = to_config(read_config(path))
    prepare_output_file(cfg)
    to_store = build_representatives(cfg, background=True)
    is_real = I.is_operator_real(cfg.hamiltonian.obj)
    dtype = with_datatype(is_real, cfg.datatype)
    # a checkpoint is read before the representatives' store starts: the
    # store then overlaps the solve, and nothing else touches the file
    # until it is done
    resume = (read_checkpoint(cfg, I.get_number_states(cfg.basis))
              if is_real and cfg.num_evals == 1 else None)
    pending = (_PendingWrite(cfg.output, BASIS_PATH + "/representatives", to_store)
               if to_store is not None else None)
    wait_io = pending.wait if pending is not None else (lambda: None)
    try:
        evals, evecs, rnorms = diagonalize(cfg, dtype, checkpoint_every=checkpoint_every,
                                           wait_io=wait_io, resume=resume)
    finally:
        wait_io()
    compute_expectations(cfg, evecs)
    return dict(evals=evals, rnorms=rnorms, seconds=time.perf_counter() - t0,
                dimension=I.get_number_states(cfg.basis), output=cfg.output,
                **diagonalize.last)


def main(argv=None) -> int:
    """`spin-ed [--debug] input_file` (app/Main.hs:31-50)."""
    p = argparse.ArgumentParser(
        prog="spin-ed",
        description="Exact diagonalization of spin systems on an AMD Instinct GPU. Takes a YAML "
                    "file describing the basis, Hamiltonian and observables.")
    p.add_argument("--debug", action="store_true", help="Enable debug output from libspined")
    p.add_argument("--checkpoint-every", type=int, default=DEFAULT_CHECKPOINT_EVERY,
                   help="Write the Lanczos checkpoint to /_workspace/lanczos every N restarts "
                        "(0: never); a later run on the same output file resumes from it")
    p.add_argument("-V", "--version", action="version", version="spined 0.1")
    p.add_argument("input_file", help="Input yaml file")
    a = p.parse_args(argv)
    logging.basicConfig(level=logging.DEBUG if a.debug else logging.INFO,
                        format="[%(asctime)s] %(levelname)s %(message)s")
    out = run_app(a.input_file, a.debug, a.checkpoint_every)
    log.info("done in %.3f s", out["seconds"])
    return 0


if __name__ == "__main__":
    sys.exit(main())
