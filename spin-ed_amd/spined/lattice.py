"""Generators and bond lists of the lattices used by the example decks and the
benchmarks (rings and periodic square lattices), plus the Heisenberg / XXZ
bond matrices.  The YAML decks of the reference (example/*.yaml) spell out the
same permutations explicitly."""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

HEISENBERG = [[1, 0, 0, 0], [0, -1, 2, 0], [0, 2, -1, 0], [0, 0, 0, 1]]


def xxz_matrix(delta: float, j: float = 1.0) -> List[List[float]]:
    d = float(delta)
    return (j * np.array([[d, 0, 0, 0], [0, -d, 2, 0], [0, 2, -d, 0], [0, 0, 0, d]])).tolist()


def ring_translation(n: int) -> List[int]:
    return [(i + 1) % n for i in range(n)]


def ring_reflection(n: int) -> List[int]:
    return [n - 1 - i for i in range(n)]


def ring_bonds(n: int) -> List[List[int]]:
    return [[i, (i + 1) % n] for i in range(n)]


def ring_symmetries(n: int, momentum: int = 0, parity: int | None = 0) -> List[Tuple[List[int], int]]:
    """[(translation, k)] plus the reflection with sector 0 (+1) or 1 (-1)."""
    gens = [(ring_translation(n), momentum)]
    if parity is not None:
        gens.append((ring_reflection(n), parity))
    return gens


def square_symmetries(L: int) -> List[Tuple[List[int], int]]:
    """Translations, both reflections and the 90-degree rotation of an L x L
    torus (site = L*y + x), all in the trivial sector (as
    example/heisenberg_square_6x6.yaml spells out for L = 6)."""
    idx = lambda x, y: (y % L) * L + (x % L)
    tx = [idx(x + 1, y) for y in range(L) for x in range(L)]
    ty = [idx(x, y + 1) for y in range(L) for x in range(L)]
    rx = [idx(L - 1 - x, y) for y in range(L) for x in range(L)]
    ry = [idx(x, L - 1 - y) for y in range(L) for x in range(L)]
    rot = [idx(L - 1 - y, x) for y in range(L) for x in range(L)]
    return [(tx, 0), (ty, 0), (rx, 0), (ry, 0), (rot, 0)]


def square_bonds(L: int) -> List[List[int]]:
    idx = lambda x, y: (y % L) * L + (x % L)
    bonds = []
    for y in range(L):
        for x in range(L):
            bonds.append([idx(x, y), idx(x + 1, y)])
            bonds.append([idx(x, y), idx(x, y + 1)])
    return bonds
