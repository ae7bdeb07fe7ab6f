"""The batched-solver workload at full size: 512 independent XXZ rings
(spined.batch.xxz_instances(512): N uniform over {12, ..., 26}, Delta ~ U(0, 2),
J_b ~ U(0.8, 1.2), Hamming weight N/2, spin inversion +1; 4.8e8 rows in
all) solved in one batched call.

  * every instance converges to residual <= 1e-9 (status 0);
  * a sample of 17 instances -- two of every size and three of N = 26 --
    matches the CPU oracle's ground energies (tests/golden/xxz512_e0.npz,
    written by tests/golden/make_xxz512_golden.py; an oracle solve takes
    minutes at N = 26) to 1e-10 relative;
  * the cases with closed forms, only for N/2 even (where the inversion-even
    sector holds the ground state), at every such size up to 24: Delta = 0
    with random J (XX chain = free fermions) and Delta = 1 with J = 1 (Bethe
    ansatz).
"""
from pathlib import Path

import numpy as np
import pytest

from spined import batch
from test_oracle_pins import bethe_ground_energy, xx_free_fermion_energy

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "xxz512_e0.npz"


def test_xxz512_all_instances_converge_and_match_oracle():
    inst = batch.xxz_instances(512)
    sysb = batch.BatchSystem(inst)
    dims = sysb.dimensions
    assert int(dims.sum()) == 481352358  # sum of C(N, N/2) / 2
    r = batch.batch_eigsh(sysb.ops, tol=1e-9)
    print(r["stats"], flush=True)
    assert r["status"] == 0
    assert int(r["stats"]["converged"]) == 512
    assert np.all(r["rnorms"] <= 1e-9), np.max(r["rnorms"])
    g = np.load(GOLDEN, allow_pickle=False)
    sizes = set()
    for i, n, delta, e0 in zip(g["index"], g["n"], g["delta"], g["e0"]):
        x = inst[int(i)]
        assert x["n"] == n and x["delta"] == delta, "workload generator changed"
        assert abs(r["evals"][i] - e0) <= 1e-10 * abs(e0), (int(i), int(n), r["evals"][i], e0)
        sizes.add(int(n))
    assert {22, 24, 26} <= sizes


def test_xxz_batch_closed_forms_every_even_half_size():
    rng = np.random.default_rng(11)
    inst = []
    for n in (12, 16, 20, 24):
        inst.append({"n": n, "delta": 0.0, "J": rng.uniform(0.8, 1.2, n).tolist()})
        inst.append({"n": n, "delta": 1.0, "J": [1.0] * n})
    sysb = batch.BatchSystem(inst)
    r = batch.batch_eigsh(sysb.ops, tol=1e-9)
    assert r["status"] == 0
    assert np.all(r["rnorms"] <= 1e-9)
    for i, x in enumerate(inst):
        exact = xx_free_fermion_energy(np.array(x["J"])) if x["delta"] == 0.0 \
            else bethe_ground_energy(x["n"])
        assert abs(r["evals"][i] - exact) <= 1e-10 * abs(exact), (x["n"], x["delta"])
