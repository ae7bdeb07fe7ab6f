"""Writes tests/golden/xxz512_e0.npz: the CPU oracle's ground energies of a
sample of the 512-instance batched XXZ workload.

The workload is spined.batch.xxz_instances(512): N uniform over
{12, 14, ..., 26}, Delta ~ U(0, 2) per instance, J_b ~ U(0.8, 1.2) per bond
(b, b+1 mod N), from numpy.random.default_rng(20261202); Hamming weight N/2
and spin inversion +1.  The sample is the first two instances of every size
and the first three of N = 26 (the size that holds three quarters of the
batch's rows).  An oracle solve takes minutes at N = 26 on a few CPU cores,
too long for a GPU test, so the energies are computed here once.

Arrays (numpy.load with allow_pickle=False):
  index   instance numbers in xxz_instances(512)
  n       their sizes
  delta   their anisotropies (a check that the workload generator is unchanged)
  e0      oracle ground energies (the oracle's Lanczos, residual <= 1e-10)

Run from the repository root:  python tests/golden/make_xxz512_golden.py
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "oracle"), str(ROOT / "spin-ed_amd")]
import pyoracle  # noqa: E402
from spined import batch, lattice  # noqa: E402

OUT = Path(__file__).resolve().parent / "xxz512_e0.npz"


def sample(inst):
    picks = []
    for n in (12, 14, 16, 18, 20, 22, 24, 26):
        idx = [i for i, x in enumerate(inst) if x["n"] == n]
        picks += idx[: 3 if n == 26 else 2]
    return picks


def oracle_e0(x, tol=1e-10):
    n = x["n"]
    o = pyoracle.System(n, n // 2, 1, [])
    for k in range(n):
        o.add_interaction(np.asarray(lattice.xxz_matrix(x["delta"], x["J"][k]),
                                     dtype=np.complex128), [[k, (k + 1) % n]])
    o.build()
    r = o.lanczos(k=1, tol=tol, max_iter=400)
    assert r["rnorms"][0] <= tol, r["rnorms"]
    return float(r["evals"][0])


def main():
    inst = batch.xxz_instances(512)
    picks = sample(inst)
    e0 = []
    for i in picks:
        t = time.perf_counter()
        e0.append(oracle_e0(inst[i]))
        print(f"instance {i} N={inst[i]['n']} E0={e0[-1]:.15f} ({time.perf_counter() - t:.1f} s)",
              flush=True)
    np.savez(OUT, index=np.array(picks, dtype=np.int64),
             n=np.array([inst[i]["n"] for i in picks], dtype=np.int64),
             delta=np.array([inst[i]["delta"] for i in picks]), e0=np.array(e0))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
