"""Pooled eigensolver vectors (blocks of 256 MiB or more are kept after a
solve and reused): the FP32 phase of a mixed solve hands its FP64-sized
blocks to the FP64 phase, a second solve of the same size takes the blocks
the first one left behind with their old contents, and releasing the pool
(spined_device_release_pool) returns them to the driver.  The solver is
deterministic, so every solve gives the same bits.  A dropped gather cache
leaves its entry block in the pool, and the next build of the same
operator's cache overwrites it in place."""
import numpy as np
import pytest

import spined
from helpers import engine, ring

pytestmark = pytest.mark.gpu


def test_pool_reuse_and_release():
    spec = ring(36, k=0, p=0, z=1)   # D = 63,068,876: 505 MB per FP64 vector
    basis, op = engine(spec)
    assert spined.get_number_states(basis) * 8 >= 256 << 20
    runs = []
    for precision, release in [("mixed", False), ("mixed", False), ("fp64", False),
                               ("mixed", True), ("fp64", False)]:
        if release:
            spined.release_device_pool()
        r = spined.eigsh(op, k=1, tol=1e-10, precision=precision, want_vectors=False)
        print(precision, r.evals, r.rnorms, r.stats["matvecs"], flush=True)
        assert r.status == 0 and r.rnorms[0] <= 1e-10
        runs.append((precision, r.evals[0], r.rnorms[0], r.stats["matvecs"]))
    mixed = [x for x in runs if x[0] == "mixed"]
    fp64 = [x for x in runs if x[0] == "fp64"]
    assert all(x[1:] == mixed[0][1:] for x in mixed)
    assert all(x[1:] == fp64[0][1:] for x in fp64)
    assert abs(mixed[0][1] - fp64[0][1]) <= 1e-10 * abs(fp64[0][1])
    spined.release_device_pool()


def test_cache_rebuild_in_pooled_block():
    spec = ring(36, k=0, p=0, z=1)   # about 4.5 GB of cache entries
    basis, op = engine(spec)
    D = spined.get_number_states(basis)
    x = np.random.default_rng(11).uniform(-1, 1, D)
    spined.set_gather_cache(op, 0)
    y0 = spined.apply(op, x)          # orbit and index kernels
    for release in (False, False, True):
        if release:
            spined.release_device_pool()
        spined.set_gather_cache(op, 1)
        y = spined.apply(op, x)       # cache (re)built, from the pool when not released
        assert spined.gather_cache(op) == 1
        np.testing.assert_array_equal(y, y0)
        spined.set_gather_cache(op, 0)  # entries back to the pool
    spined.release_device_pool()
