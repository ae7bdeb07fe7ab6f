"""Batched uneven solves (spined_batch_eigsh): many XXZ rings solved in one
persistent launch.

The synthetic workload is 512 independent XXZ rings: N uniform over
{12, 14, ..., 26}, anisotropy Delta ~ U(0, 2) per instance, bond couplings
J_b ~ U(0.8, 1.2) on bonds (b, b+1 mod N), drawn in that order from
numpy.random.default_rng(20261202).  Translation is broken, so only the
Hamming weight N/2 and spin inversion +1 apply.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List

import numpy as np

from . import _lib
from . import internal as I
from . import lattice


def xxz_instances(count: int = 512, seed: int = 20261202,
                  sizes=(12, 14, 16, 18, 20, 22, 24, 26)) -> List[dict]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(sizes))
        delta = float(rng.uniform(0.0, 2.0))
        J = rng.uniform(0.8, 1.2, n).tolist()
        out.append({"n": n, "delta": delta, "J": J})
    return out


class BatchSystem:
    """Bases (one per size, shared) and one operator per instance."""

    def __init__(self, instances: List[dict]):
        self.instances = instances
        self.bases: Dict[int, I.SpinBasis] = {}
        self.ops: List[I.Operator_] = []
        for inst in instances:
            n = inst["n"]
            if n not in self.bases:
                b = I.mk_basis(I.mk_group([]), n, n // 2, 1)
                I.build_basis(b)
                self.bases[n] = b
            terms = [I.mk_interaction(lattice.xxz_matrix(inst["delta"], inst["J"][k]),
                                      [[k, (k + 1) % n]]) for k in range(n)]
            self.ops.append(I.mk_operator(self.bases[n], terms))

    @property
    def dimensions(self) -> np.ndarray:
        return np.array([I.get_number_states(self.bases[i["n"]]) for i in self.instances])


class SpinedBatchStats(C.Structure):
    _fields_ = [("cycles", C.c_int32), ("converged", C.c_int32), ("seconds_total", C.c_double),
                ("seconds_cycles", C.c_double), ("total_rows", C.c_uint64),
                ("matvecs_per_instance_max", C.c_int64)]


def batch_eigsh(ops: List[I.Operator_], tol: float = 1e-9, max_basis_size: int = 0,
                seed: int = 0x5EED, max_cycles: int = 200) -> dict:
    """One persistent launch for all instances (spined_batch_eigsh); max_basis_size
    0 takes the library default (16)."""
    L = _lib.load()
    f = L.spined_batch_eigsh
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.c_int32, C.c_uint64, C.c_int32,
                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.POINTER(SpinedBatchStats)]
    n = len(ops)
    arr = (C.c_void_p * max(n, 1))(*[o.ptr.value for o in ops])
    ev = np.zeros(n)
    rn = np.zeros(n)
    cy = np.zeros(n, dtype=np.int32)
    sec = np.zeros(n)
    mv = np.zeros(n, dtype=np.int64)
    st = SpinedBatchStats()
    status = f(arr, n, tol, max_basis_size, seed, max_cycles, ev.ctypes.data, rn.ctypes.data,
               cy.ctypes.data, sec.ctypes.data, mv.ctypes.data, C.byref(st))
    if status not in (0, 10):
        I.check_status(status)
    return {"status": status, "evals": ev, "rnorms": rn, "cycles": cy, "seconds": sec,
            "matvecs": mv, "stats": {k: getattr(st, k) for k, _ in SpinedBatchStats._fields_}}
