"""GPU checks of the batched solver against one-at-a-time oracle solves and
closed forms (XX free fermions, Bethe ansatz) for the special cases."""
import numpy as np
import pytest

import pyoracle
from spined import batch, lattice
from test_oracle_pins import bethe_ground_energy, xx_free_fermion_energy

pytestmark = pytest.mark.gpu


def oracle_e0(inst):
    n = inst["n"]
    o = pyoracle.System(n, n // 2, 1, [])
    for k in range(n):
        o.add_interaction(np.asarray(lattice.xxz_matrix(inst["delta"], inst["J"][k]),
                                     dtype=np.complex128), [[k, (k + 1) % n]])
    o.build()
    return o.lanczos(k=1, tol=1e-11, max_iter=400)["evals"][0]


def test_batch_matches_oracle_per_instance():
    inst = batch.xxz_instances(24, seed=7, sizes=(8, 10, 12, 14))
    sysb = batch.BatchSystem(inst)
    r = batch.batch_eigsh(sysb.ops, tol=1e-9, max_basis_size=16)
    print(r["stats"], r["cycles"])
    assert r["status"] == 0
    assert np.all(r["rnorms"] <= 1e-9)
    for i, x in enumerate(inst):
        e = oracle_e0(x)
        assert abs(r["evals"][i] - e) <= 1e-9 * abs(e), (i, x["n"], r["evals"][i], e)


def test_batch_special_cases_closed_forms():
    # Delta = 0 (XX, random J) -> free fermions; Delta = 1, J = 1 -> Bethe ansatz;
    # both are the inversion-even ground state only for N/2 even
    rng = np.random.default_rng(3)
    inst = []
    for n in (12, 16, 20):
        inst.append({"n": n, "delta": 0.0, "J": rng.uniform(0.8, 1.2, n).tolist()})
        inst.append({"n": n, "delta": 1.0, "J": [1.0] * n})
    sysb = batch.BatchSystem(inst)
    r = batch.batch_eigsh(sysb.ops, tol=1e-9)
    assert r["status"] == 0
    for i, x in enumerate(inst):
        exact = xx_free_fermion_energy(np.array(x["J"])) if x["delta"] == 0.0 \
            else bethe_ground_energy(x["n"])
        assert abs(r["evals"][i] - exact) <= 1e-9 * abs(exact), (x["n"], x["delta"])
