"""Learned start vector: training data from the engine's own ground states,
GPU training of the periodic CNN (BF16 MFMA forward/backward in cnn.hip) and
evaluation over all representatives of a target basis.

Physics conventions (ring of N sites, bonds (i, i+1)):
  * Marshall prior: sign psi(s) = (-1)^{#up spins on even sites} x global
    sign.  Translations, reflections and the flip multiply it by (-1)^{N/2},
    so the ground state lies in (k=0, p=+, z=+) when N/2 is even and in
    (k=pi, p=-, z=-) when N/2 is odd.  The network learns log|psi| and the
    residual sign after the Marshall rotation (+1 for the Heisenberg ring).
  * In the symmetric basis psi(r) = c_r N_r and a start vector is
    c_r = psi(r) / N_r, N_r = sqrt(|Stab r| / |G|).
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np

from . import _lib
from . import internal as I
from . import lattice
from .device import DeviceArray
from .solver import eigsh

NPARAMS = 2722
CH, KW = 16, 5


def marshall_mask(n: int) -> int:
    return sum(1 << i for i in range(0, n, 2))


def ring_sector_for_ground_state(n: int):
    """(k, p_sector, z) of the ring's ground state (Marshall rule)."""
    if (n // 2) % 2 == 0:
        return 0, 0, 1
    return n // 2, 1, -1


def init_params(seed: int = 1) -> np.ndarray:
    """He-style initialisation (fixed scheme, seeded)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, dtype=np.float32)
    o = 0
    p[o:o + CH * KW] = rng.normal(0, math.sqrt(2.0 / KW), CH * KW); o += CH * KW
    o += CH
    p[o:o + CH * CH * KW] = rng.normal(0, math.sqrt(2.0 / (CH * KW)), CH * CH * KW); o += CH * CH * KW
    o += CH
    p[o:o + CH * CH * KW] = rng.normal(0, math.sqrt(2.0 / (CH * KW)), CH * CH * KW); o += CH * CH * KW
    o += CH
    p[o:o + CH] = rng.normal(0, 0.01, CH); o += CH   # u
    o += 1                                           # u0
    p[o:o + CH] = rng.normal(0, 0.01, CH); o += CH   # v
    p[o] = 2.0                                       # v0: positive residual sign prior
    return p


class CNN:
    """Handle on a spined_cnn model (FP32 master weights on the GPU)."""

    def __init__(self, params: Optional[np.ndarray] = None, seed: int = 1):
        L = _lib.load()
        assert L.spined_cnn_num_params() == NPARAMS
        p = init_params(seed) if params is None else np.ascontiguousarray(params, dtype=np.float32)
        h = C.c_void_p()
        I.check_status(L.spined_cnn_create(C.byref(h), p.ctypes.data))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().spined_cnn_destroy(self._h)
            self._h = C.c_void_p(None)

    @property
    def handle(self):
        return self._h

    def params(self) -> np.ndarray:
        out = np.empty(NPARAMS, dtype=np.float32)
        I.check_status(_lib.load().spined_cnn_get_params(self._h, out.ctypes.data))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        I.check_status(_lib.load().spined_cnn_set_params(self._h, p.ctypes.data))

    def forward(self, n: int, states: np.ndarray, store: bool = False):
        st = DeviceArray.from_numpy(np.ascontiguousarray(states, dtype=np.uint64))
        cnt = len(states)
        a = DeviceArray(cnt, np.float32)
        t = DeviceArray(cnt, np.float32)
        z = DeviceArray(3 * ((cnt + 15) // 16 * 16) * n * CH, np.float32) if store else None
        I.check_status(_lib.load().spined_cnn_forward_device(
            self._h, n, cnt, st.ptr, a.ptr, t.ptr, z.ptr if z else None))
        if store:
            return a, t, z, st
        return a.numpy(), t.numpy()

    def gradient(self, n: int, states: np.ndarray, ga: np.ndarray, gt: np.ndarray) -> np.ndarray:
        """d/dparams of sum ga*a + gt*t (for the check against the oracle)."""
        a, t, z, st = self.forward(n, states, store=True)
        dga = DeviceArray.from_numpy(np.ascontiguousarray(ga, dtype=np.float32))
        dgt = DeviceArray.from_numpy(np.ascontiguousarray(gt, dtype=np.float32))
        g = DeviceArray(NPARAMS, np.float32)
        I.check_status(_lib.load().spined_cnn_backward_device(
            self._h, n, len(states), st.ptr, z.ptr, dga.ptr, dgt.ptr, g.ptr))
        return g.numpy()

    def train_step(self, n: int, states: np.ndarray, target_logamp: np.ndarray,
                   target_sign: np.ndarray, lr: float) -> float:
        """One Adam step on a batch of one system size; the log-amplitude loss
        is invariant to a per-size constant (the network's output is
        unnormalised)."""
        a_d, t_d, z, st = self.forward(n, states, store=True)
        a, t = a_d.numpy().astype(np.float64), t_d.numpy().astype(np.float64)
        B = len(states)
        r = a - target_logamp
        r -= r.mean()
        ga = (2.0 * r / B).astype(np.float32)
        sig = target_sign
        gt = (-sig / (1.0 + np.exp(np.clip(sig * t, -50, 50))) / B).astype(np.float32)
        dga = DeviceArray.from_numpy(ga)
        dgt = DeviceArray.from_numpy(gt)
        I.check_status(_lib.load().spined_cnn_backward_device(
            self._h, n, B, st.ptr, z.ptr, dga.ptr, dgt.ptr, None))
        I.check_status(_lib.load().spined_cnn_adam_step(self._h, C.c_float(lr)))
        return float(np.mean(r * r))

    def start_vector_device(self, basis: I.SpinBasis, n: int) -> DeviceArray:
        D = I.get_number_states(basis)
        v = DeviceArray(D, np.float64)
        I.check_status(_lib.load().spined_cnn_start_vector(self._h, basis.ptr, marshall_mask(n),
                                                           v.ptr))
        return v


@dataclass
class GroundState:
    n: int
    reps: np.ndarray
    logamp: np.ndarray   # log |psi(r)|
    sign: np.ndarray     # residual sign after the Marshall rotation (+-1)
    weight: np.ndarray   # c_r^2 (probability of the orbit)
    energy: float
    coeffs: np.ndarray   # c_r
    norm: np.ndarray     # N_r


def ring_system(n: int, k: int, p: int, z: int):
    gens = lattice.ring_symmetries(n, k, p)
    basis = I.mk_basis(I.mk_group([I.mk_symmetry(q, s) for q, s in gens]), n, n // 2, z)
    op = I.mk_operator(basis, [I.mk_interaction(lattice.HEISENBERG, lattice.ring_bonds(n))])
    return basis, op


def ground_state(n: int, tol: float = 1e-8) -> GroundState:
    """The engine's own ground state of the Heisenberg ring (training data)."""
    k, p, z = ring_sector_for_ground_state(n)
    basis, op = ring_system(n, k, p, z)
    I.build_basis(basis)
    res = eigsh(op, k=1, tol=tol)
    c = res.evecs[:, 0]
    reps = I.basis_get_states(basis)
    stab = I.basis_get_stabilizers(basis).astype(np.float64)
    G = 4 * n
    norm = np.sqrt(stab / G)
    psi = c * norm
    m = np.array([(-1) ** bin(int(x) & marshall_mask(n)).count("1") for x in reps]) \
        if len(reps) < 200000 else _marshall_signs(reps, n)
    rot = np.sign(psi) * m
    s0 = 1.0 if np.sum(rot * c * c) >= 0 else -1.0
    with np.errstate(divide="ignore"):
        logamp = np.log(np.abs(psi))
    logamp[~np.isfinite(logamp)] = -60.0
    return GroundState(n, reps, logamp, rot * s0, c * c, float(res.evals[0]), c, norm)


def _marshall_signs(reps: np.ndarray, n: int) -> np.ndarray:
    x = reps & np.uint64(marshall_mask(n))
    cnt = np.zeros(len(reps), dtype=np.int64)
    for b in range(0, n, 2):
        cnt += ((x >> np.uint64(b)) & np.uint64(1)).astype(np.int64)
    return np.where(cnt % 2 == 0, 1.0, -1.0)


def train(data: List[GroundState], steps: int = 600, batch: int = 2048, lr: float = 3e-3,
          seed: int = 7, model: Optional[CNN] = None, log=None) -> CNN:
    """Adam on batches drawn with probability c_r^2 from each size in turn."""
    rng = np.random.default_rng(seed)
    model = model or CNN(seed=seed)
    cdfs = []
    for d in data:
        c = np.cumsum(d.weight)
        cdfs.append(c / c[-1])
    t0 = time.perf_counter()
    for it in range(steps):
        j = it % len(data)
        d = data[j]
        idx = np.minimum(np.searchsorted(cdfs[j], rng.random(batch)), len(d.reps) - 1)
        cur_lr = lr * (0.5 * (1 + math.cos(math.pi * it / steps)) * 0.9 + 0.1)
        loss = model.train_step(d.n, d.reps[idx], d.logamp[idx], d.sign[idx], cur_lr)
        if log and (it % 100 == 0 or it == steps - 1):
            log(f"step {it} N={d.n} loss {loss:.4e} ({time.perf_counter() - t0:.1f} s)")
    return model


def start_coefficients(model: CNN, gs: GroundState) -> np.ndarray:
    """Host-side c_r = psi_model(r) / N_r on a solved size (normalised)."""
    a, t = model.forward(gs.n, gs.reps)
    m = _marshall_signs(gs.reps, gs.n)
    amp = np.exp(a.astype(np.float64) - a.max()) * np.where(t < 0, -1.0, 1.0) * m
    c = amp / gs.norm
    return c / np.linalg.norm(c)


def overlap(model: CNN, gs: GroundState) -> float:
    """|<start|psi0>| of the learned start vector on a solved size."""
    return float(abs(np.dot(start_coefficients(model, gs), gs.coeffs)))


def save(model: CNN, path: str) -> None:
    np.save(path, model.params())


def load(path: str) -> CNN:
    return CNN(params=np.load(path, allow_pickle=False))


def model_for(cfg, model=None) -> CNN:
    """The network of the YAML front end (`start_vector: learned`,
    `model: params.npy`), or the one given."""
    if model is None:
        if not cfg.spec.model:
            raise ValueError("start_vector: learned needs `model:` (a .npy parameter file)")
        model = load(cfg.spec.model)
    return model


def start_vector_for(cfg, model=None):
    """Host copy of the learned start vector of the YAML front end; the
    ring's Marshall sublattice is the even sites."""
    n = cfg.spec.basis.number_spins
    v = model_for(cfg, model).start_vector_device(cfg.basis, n)
    return v.numpy()
