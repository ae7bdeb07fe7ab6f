"""Shared construction of the same spin system in the engine (libspined via
the spined package) and in the CPU oracle."""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

import pyoracle
import spined
from spined import lattice


@dataclass
class Spec:
    n: int
    w: Optional[int]
    z: Optional[int]
    gens: List[Tuple[List[int], int]]
    terms: List[Tuple[list, list]] = field(default_factory=list)  # (matrix, sites)
    name: str = ""


def ring(n, k=0, p=0, z=1, w="half", matrix=None, name=None):
    gens = [] if k is None else lattice.ring_symmetries(n, k, p)
    ww = n // 2 if w == "half" else w
    m = lattice.HEISENBERG if matrix is None else matrix
    return Spec(n, ww, z, gens, [(m, lattice.ring_bonds(n))], name or f"ring{n}_k{k}_p{p}_z{z}")


def torus_symmetries(l1, l2, kx=0, ky=0, reflect=False):
    """Translations of a row-major l1 x l2 torus (site = l1 y + x) in
    momentum sectors (kx, ky), optionally with the in-row reflection
    (sector 0)."""
    idx = lambda x, y: (y % l2) * l1 + (x % l1)
    gens = [([idx(x + 1, y) for y in range(l2) for x in range(l1)], kx),
            ([idx(x, y + 1) for y in range(l2) for x in range(l1)], ky)]
    if reflect:
        gens.append(([idx(l1 - 1 - x, y) for y in range(l2) for x in range(l1)], 0))
    return gens


def torus_bonds(l1, l2):
    idx = lambda x, y: (y % l2) * l1 + (x % l1)
    return [[idx(x, y), idx(x + 1, y)] for y in range(l2) for x in range(l1)] + \
           [[idx(x, y), idx(x, y + 1)] for y in range(l2) for x in range(l1)]


def torus(l1, l2, kx=0, ky=0, reflect=False, z=None, w="half", name=None):
    n = l1 * l2
    return Spec(n, n // 2 if w == "half" else w, z, torus_symmetries(l1, l2, kx, ky, reflect),
                [(lattice.HEISENBERG, torus_bonds(l1, l2))],
                name or f"torus{l1}x{l2}_k{kx}{ky}_r{int(reflect)}_z{z}")


def engine(spec: Spec, build=True):
    syms = [spined.mk_symmetry(p, s) for p, s in spec.gens]
    basis = spined.mk_basis(spined.mk_group(syms), spec.n, spec.w, spec.z)
    if build:
        spined.build_basis(basis)
    terms = [spined.mk_interaction(m, s) for m, s in spec.terms]
    op = spined.mk_operator(basis, terms)
    return basis, op


def oracle(spec: Spec, build=True):
    o = pyoracle.System(spec.n, spec.w, spec.z if spec.z else None, spec.gens)
    for m, s in spec.terms:
        o.add_interaction(np.asarray(m, dtype=np.complex128), s)
    if build:
        o.build()
    return o
