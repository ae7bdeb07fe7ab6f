"""YAML configuration -- mirror of the FromJSON instances of src/SpinED.hs:81-173.

Keys, optionality and defaults follow the reference parser (which is
authoritative where template.yaml disagrees):
  basis: {number_spins, hamming_weight?, spin_inversion?, symmetries: [{permutation, sector}]}
  hamiltonian: {name, terms: [{matrix, sites}]}
  observables: [operator]
  output = "exact_diagonalization_result.h5", number_vectors = 1, precision = 0.0,
  datatype = float64 (float32 | float64, case-insensitive),
  max_primme_basis_size, max_primme_block_size, min_primme_restart_size
  (defaults: the engine's automatic choices, standing in for primmeDefaults).
Engine-only optional keys (old decks stay valid): solver_precision
("fp64" | "mixed"), start_vector ("random" | "learned"), seed, model.
Complex numbers are a number or a [re, im] pair (FromJSON (Complex Double),
SpinED.hs:114-122).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import yaml


class ConfigError(ValueError):
    pass


def _complex(v):
    if isinstance(v, bool):
        raise ConfigError(f"expected a number, got {v!r}")
    if isinstance(v, (int, float)):
        return complex(float(v), 0.0)
    if isinstance(v, (list, tuple)) and len(v) == 2:
        return complex(float(v[0]), float(v[1]))
    raise ConfigError(f"expected a complex number (number or [re, im]), got {v!r}")


@dataclass
class SymmetrySpec:
    permutation: List[int]
    sector: int


@dataclass
class BasisSpec:
    number_spins: int
    hamming_weight: Optional[int]
    spin_inversion: Optional[int]
    symmetries: List[SymmetrySpec]


@dataclass
class InteractionSpec:
    matrix: List[List[complex]]
    sites: List[List[int]]


@dataclass
class OperatorSpec:
    name: str
    terms: List[InteractionSpec]


@dataclass
class ConfigSpec:
    basis: BasisSpec
    hamiltonian: OperatorSpec
    observables: List[OperatorSpec]
    output: str = "exact_diagonalization_result.h5"
    number_vectors: int = 1
    precision: float = 0.0
    datatype: str = "float64"
    max_primme_basis_size: int = 0
    max_primme_block_size: int = 0
    min_primme_restart_size: int = 0
    solver_precision: str = "fp64"
    start_vector: str = "random"
    seed: int = 0x5EED
    model: Optional[str] = None


def _req(d, key, what):
    if not isinstance(d, dict):
        raise ConfigError(f"parsing {what} failed, expected Object")
    if key not in d:
        raise ConfigError(f"key \"{key}\" not found in {what}")
    return d[key]


def parse_symmetry(d) -> SymmetrySpec:
    perm = _req(d, "permutation", "symmetry")
    sec = _req(d, "sector", "symmetry")
    if not isinstance(perm, list) or not all(isinstance(p, int) for p in perm):
        raise ConfigError("permutation must be a list of integers")
    if not isinstance(sec, int):
        raise ConfigError("sector must be an integer")
    return SymmetrySpec(list(perm), int(sec))


def parse_basis(d) -> BasisSpec:
    n = _req(d, "number_spins", "basis")
    syms = _req(d, "symmetries", "basis")
    if not isinstance(syms, list):
        raise ConfigError("symmetries must be a list")
    hw = d.get("hamming_weight")
    si = d.get("spin_inversion")
    return BasisSpec(int(n), None if hw is None else int(hw), None if si is None else int(si),
                     [parse_symmetry(s) for s in syms])


def parse_interaction(d) -> InteractionSpec:
    m = _req(d, "matrix", "interaction")
    s = _req(d, "sites", "interaction")
    if not isinstance(m, list) or not all(isinstance(r, list) for r in m):
        raise ConfigError("matrix must be a list of rows")
    return InteractionSpec([[_complex(v) for v in row] for row in m], [list(map(int, t)) for t in s])


def parse_operator(d) -> OperatorSpec:
    name = _req(d, "name", "operator")
    terms = _req(d, "terms", "operator")
    return OperatorSpec(str(name), [parse_interaction(t) for t in terms])


def parse_datatype(v) -> str:
    s = str(v).lower()
    if s not in ("float32", "float64"):
        raise ConfigError(
            f"parsing Datatype failed, expected either 'float32' or 'float64', but got '{v}'")
    return s


def parse_config(d) -> ConfigSpec:
    """FromJSON ConfigSpec (SpinED.hs:161-173)."""
    c = ConfigSpec(
        basis=parse_basis(_req(d, "basis", "config")),
        hamiltonian=parse_operator(_req(d, "hamiltonian", "config")),
        observables=[parse_operator(o) for o in (_req(d, "observables", "config") or [])],
    )
    if "output" in d:
        c.output = str(d["output"])
    if "number_vectors" in d:
        c.number_vectors = int(d["number_vectors"])
    if "precision" in d:
        c.precision = float(d["precision"])
    if "datatype" in d:
        c.datatype = parse_datatype(d["datatype"])
    for key in ("max_primme_basis_size", "max_primme_block_size", "min_primme_restart_size"):
        if key in d:
            setattr(c, key, int(d[key]))
    if "solver_precision" in d:
        c.solver_precision = str(d["solver_precision"]).lower()
        if c.solver_precision not in ("fp64", "mixed"):
            raise ConfigError("solver_precision must be 'fp64' or 'mixed'")
    if "start_vector" in d:
        c.start_vector = str(d["start_vector"]).lower()
        if c.start_vector not in ("random", "learned"):
            raise ConfigError("start_vector must be 'random' or 'learned'")
    if "seed" in d:
        c.seed = int(d["seed"])
    if "model" in d:
        c.model = str(d["model"])
    return c


def read_config(path: str) -> ConfigSpec:
    """readConfig (SpinED.hs:250-258)."""
    with open(path) as f:
        data = yaml.safe_load(f)
    return parse_config(data)


def loads(text: str) -> ConfigSpec:
    return parse_config(yaml.safe_load(text))
