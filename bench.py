"""Benchmark: seconds to E0 (residual <= 1e-10) of the Heisenberg ring on one
MI355X, with the matvec's HBM roofline and a CPU baseline.

A step is one full ground-state solve (basis already built and resident in
HBM): start vector -> Lanczos on the GPU -> explicit residual <= tol.
`value` = mean seconds per solve over the K timed steps (lower is better).

roofline: the matvec kernel (dominant kernel), timed with HIP events that
libspined records on its own stream around every matvec launch.  Algorithmic
bytes per matvec: B = D (8 + 4 + s_x + s_y) + E_off s_x  (8: representative,
4: stored stabiliser size, s_x / s_y: bytes of one x / y element, E_off: the
off-diagonal terms that gather x, counted on the GPU).

cpu_baseline: the oracle (plain C + OpenMP on all host cores) timed on a
bounded sample of rows of the same matvec, extrapolated to the solve with the
GPU run's matvec count ("port", extrapolated).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "spin-ed_amd"))

import numpy as np  # noqa: E402

CONFIGS = {
    # name: (N, description)
    "ring40": (40, "Heisenberg ring N=40, w=20, k=0, p=+1, z=+1 (861,725,794 states)"),
    "ring36": (36, "Heisenberg ring N=36, w=18, k=0, p=+1, z=+1"),
    "ring32": (32, "Heisenberg ring N=32, w=16, k=0, p=+1, z=+1"),
    "ring28": (28, "Heisenberg ring N=28, w=14, k=0, p=+1, z=+1"),
    "ring24": (24, "Heisenberg ring N=24, w=12, k=0, p=+1, z=+1 (28,968 states)"),
    "square6x6": (36, "Heisenberg 6x6 square lattice, w=18, 36 translations x D4 x spin "
                      "inversion: K=8 block Lanczos (FP64, FP32 storage) and one-vector FP64"),
    "xxz512": (0, "512 XXZ rings, N uniform over {12..26}, Delta~U(0,2), J_b~U(0.8,1.2), "
                  "w=N/2, z=+1, residual 1e-9, one persistent batched solver launch"),
}
PEAK_HBM = 8.0e12


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None)
    p.add_argument("--warmup", type=int, default=None)
    p.add_argument("--config", default=os.environ.get("SPINED_BENCH_CONFIG", "ring40"),
                   choices=sorted(CONFIGS))
    p.add_argument("--precision", default="mixed", choices=["fp64", "mixed"],
                   help="headline storage mode (mixed = FP32-storage phase first)")
    p.add_argument("--start", default="learned", choices=["random", "learned"])
    p.add_argument("--train-sizes", default="16,18,20,22,24,26,28,30,32")
    p.add_argument("--train-steps", type=int, default=2400)
    p.add_argument("--no-compare", action="store_true",
                   help="skip the plain FP64 random-start reference solve")
    p.add_argument("--tol", type=float, default=1e-10)
    p.add_argument("--basis-size", type=int, default=0)
    p.add_argument("--cpu-sample-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-square", action="store_true",
                   help="skip the 6x6 square-lattice block Lanczos line (block_square6x6)")
    p.add_argument("--no-batch", action="store_true",
                   help="skip the 512-instance batched solve reported beside the headline")
    p.add_argument("--traffic-json", default=None,
                   help="JSON with measured FETCH/WRITE bytes per matvec launch")
    p.add_argument("--copy-calibration-bytes", type=int, default=0,
                   help="run the plain copy kernel once on this many bytes (counter calibration)")
    p.add_argument("--gather-calibration-lines", type=int, default=0,
                   help="run the random-gather calibration kernel over this many 128-byte lines "
                        "(a power of two), once reading 4 bytes and once the whole line of each")
    p.add_argument("--start-vector-only", type=int, default=0,
                   help="profiling mode: evaluate the CNN start vector over every representative "
                        "this many times (untrained parameters) instead of solves")
    p.add_argument("--matvecs-only", type=int, default=0,
                   help="profiling mode: run this many FP64 matvecs instead of solves")
    p.add_argument("--block", type=int, default=1, choices=(1, 2, 4, 8),
                   help="matvecs-only mode: row-interleaved block width")
    p.add_argument("--index", default="auto", choices=("auto", "bucket", "hash"),
                   help="index of representatives (spined_basis_set_index_kind)")
    p.add_argument("--gather-cache", choices=["auto", "on", "off"], default="auto",
                   help="gather cache of the operator (auto: built when it fits in free memory)")
    p.add_argument("--no-autotune", action="store_true",
                   help="skip the matvec autotuner (run before the timed region)")
    p.add_argument("--bucket-bits", type=int, default=-1,
                   help="bucket table bits of the bucket index (-1: automatic)")
    p.add_argument("--solve-block", type=int, default=0, choices=(0, 2, 4, 8),
                   help="block width of the solver's first phase (max_block_size); 0 = one-vector")
    p.add_argument("--mv-precision", default="fp64", choices=("fp64", "mixed"),
                   help="matvecs-only mode: FP64 x, or FP32 x with FP64 y (the FP32-storage phase)")
    return p.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a design "
                                   "note only (rows by contiguous representative ranges + "
                                   "all-gather of x), not a measured path", "n_gpus": a.gpus}))
        return 2
    import spined
    from spined import lattice
    from spined.device import DeviceArray, device_synchronize

    n, desc = CONFIGS[a.config]
    if a.config == "square6x6":
        print(json.dumps(run_square6x6()), flush=True)
        return 0
    if a.config == "xxz512":
        return bench_batch(a, desc)
    big = n >= 36
    steps = a.steps if a.steps is not None else (1 if big else 5)
    # one untimed solve brings the process to its steady state, in which the
    # device pool holds the Krylov blocks: the first mapping of ~130 GB of
    # fresh device memory costs about a second of idle device; every timed
    # solve still builds its own gather cache and start vector
    warmup = a.warmup if a.warmup is not None else 1

    gens = lattice.ring_symmetries(n, 0, 0)
    t0 = time.perf_counter()
    basis = spined.mk_basis(spined.mk_group([spined.mk_symmetry(p, s) for p, s in gens]), n,
                            n // 2, 1)
    op = spined.mk_operator(basis, [spined.mk_interaction(lattice.HEISENBERG,
                                                          lattice.ring_bonds(n))])
    spined.set_index_kind(basis, {"auto": -1, "bucket": 0, "hash": 1}[a.index])
    spined.set_gather_cache(op, {"auto": -1, "on": 1, "off": 0}[a.gather_cache])
    spined.set_bucket_bits(basis, a.bucket_bits)
    spined.build_basis(basis)
    device_synchronize()
    t_build = time.perf_counter() - t0
    D = spined.get_number_states(basis)
    log(f"{a.config}: D={D} basis build {t_build:.3f} s")
    e_off = spined.count_offdiag(op)
    log(f"off-diagonal gathers per matvec: {e_off}")
    tune = None
    if (not a.no_autotune and not a.matvecs_only and not a.start_vector_only
            and a.bucket_bits < 0 and a.index == "auto"):
        t = time.perf_counter()
        tune = spined.autotune(op)
        tune["seconds"] = round(time.perf_counter() - t, 2)
        log(f"autotune: {tune}")

    if a.copy_calibration_bytes:
        src_b = DeviceArray.zeros(a.copy_calibration_bytes // 8, np.float64)
        dst_b = DeviceArray(a.copy_calibration_bytes // 8, np.float64)
        spined.check_status(spined._lib.load().spined_debug_copy_device(
            dst_b.ptr, src_b.ptr, a.copy_calibration_bytes))
        src_b.free()
        dst_b.free()

    if a.gather_calibration_lines:
        L = spined._lib.load()
        nl = a.gather_calibration_lines
        src_b = DeviceArray.zeros(nl * 32, np.uint32)
        nw = L.spined_debug_gather_out_words()
        dst_b = DeviceArray(nw, np.uint32)
        for words in (1, 32):
            spined.check_status(L.spined_debug_gather_device(dst_b.ptr, nw, src_b.ptr, nl, words))
        src_b.free()
        dst_b.free()

    if a.start_vector_only:
        from spined import learned
        model = learned.CNN(seed=1)
        v = model.start_vector_device(basis, n)
        device_synchronize()
        t = time.perf_counter()
        for _ in range(a.start_vector_only):
            v.free()
            v = model.start_vector_device(basis, n)
        device_synchronize()
        dt = (time.perf_counter() - t) / a.start_vector_only
        print(json.dumps({"mode": "start-vector-only", "config": desc, "dimension": D,
                          "evaluations": a.start_vector_only, "s_per_evaluation": dt}))
        return 0

    # one untimed matvec warms the kernels even when warmup = 0
    x = DeviceArray.from_numpy(np.ones(D))
    y = DeviceArray(D, np.float64)
    spined.matvec_device(op, x, y, 1)
    spined.synchronize()
    if a.matvecs_only and a.block > 1:
        x.free()
        y.free()
        K = a.block
        mixed = a.mv_precision == "mixed"
        xb = DeviceArray.from_numpy(np.ones(D * K, np.float32 if mixed else np.float64))
        yb = DeviceArray(D * K, np.float64)
        code = 0 if mixed else 1
        spined.matvec_rows_device(op, xb, yb, code, K, y_is_fp64=mixed)
        spined.synchronize()
        t = time.perf_counter()
        for _ in range(a.matvecs_only):
            spined.matvec_rows_device(op, xb, yb, code, K, y_is_fp64=mixed)
        spined.synchronize()
        dt = (time.perf_counter() - t) / a.matvecs_only
        print(json.dumps({"mode": "matvecs-only", "config": desc, "dimension": D, "block": K, "index": a.index,
                          "mv_precision": a.mv_precision, "launches": a.matvecs_only,
                          "ms_per_launch": dt * 1e3, "ms_per_vector": dt * 1e3 / K,
                          "offdiag_gathers_per_launch": e_off}))
        return 0
    if a.matvecs_only:
        mixed = a.mv_precision == "mixed"
        if mixed:
            x.free()
            x = DeviceArray.from_numpy(np.ones(D, np.float32))
        code = 0 if mixed else 1
        spined.matvec_device(op, x, y, code, y_is_fp64=mixed)
        spined.synchronize()
        t = time.perf_counter()
        for _ in range(a.matvecs_only):
            spined.matvec_device(op, x, y, code, y_is_fp64=mixed)
        spined.synchronize()
        dt = (time.perf_counter() - t) / a.matvecs_only
        print(json.dumps({"mode": "matvecs-only", "config": desc, "dimension": D, "index": a.index,
                          "bucket_bits": spined.bucket_bits(basis),
                          "gather_cache": spined.gather_cache(op),
                          "mv_precision": a.mv_precision, "matvecs": a.matvecs_only,
                          "ms_per_matvec": dt * 1e3, "offdiag_gathers_per_launch": e_off}))
        return 0
    x.free()
    y.free()

    model = None
    train_info = None
    if a.start == "learned":
        from spined import learned
        t = time.perf_counter()
        sizes = [int(s) for s in a.train_sizes.split(",") if s]
        data = [learned.ground_state(m) for m in sizes]
        t_data = time.perf_counter() - t
        model = learned.train(data, steps=a.train_steps, batch=2048, lr=3e-3, log=log)
        t_train = time.perf_counter() - t - t_data
        ov = {m.n: round(learned.overlap(model, m), 6) for m in data[-2:]}
        train_info = {"sizes": sizes, "steps": a.train_steps, "data_s": round(t_data, 2),
                      "train_s": round(t_train, 2), "overlap_on_training_sizes": ov}
        log(f"learned start: {train_info}")

    cache_mode = {"auto": -1, "on": 1, "off": 0}[a.gather_cache]

    def fresh_cache():
        # every timed solve builds its own gather cache inside its own time;
        # the previous one is freed before the clock starts
        spined.set_gather_cache(op, 0)
        spined.set_gather_cache(op, cache_mode)
        device_synchronize()

    def solve(precision=a.precision, use_model=True):
        return spined.eigsh(op, k=1, tol=a.tol, precision=precision, seed=0x5EED,
                            want_vectors=False, max_basis_size=a.basis_size, verbose=True,
                            max_block_size=a.solve_block if use_model else 0,
                            model=model if use_model else None,
                            sign_mask=sum(1 << i for i in range(0, n, 2)))

    for i in range(warmup):
        r = solve()
        log(f"warmup {i}: E0={r.evals[0]:.15f} res={r.rnorms[0]:.2e} {r.stats}")

    # each step is bracketed by a device synchronisation on both sides; the
    # previous step's gather cache is freed between steps, outside the clock,
    # and every step builds its own inside it
    results = []
    elapsed = 0.0
    for i in range(steps):
        fresh_cache()
        t_step = time.perf_counter()
        r = solve()
        device_synchronize()
        elapsed += time.perf_counter() - t_step
        results.append(r)
        log(f"step {i}: E0={r.evals[0]:.15f} res={r.rnorms[0]:.2e} matvecs={r.stats['matvecs']} "
            f"t={r.stats['seconds_total']:.3f}s matvec_t={r.stats['seconds_matvec']:.3f}s")
    per_step = elapsed / steps
    last = results[-1]
    cached = spined.gather_cache(op) == 1
    matvecs = sum(r.stats["matvecs"] for r in results)
    mv_seconds = sum(r.stats["seconds_matvec"] for r in results)
    mv_avg = mv_seconds / max(matvecs, 1)

    # launches by x type: FP32 x (the FP32-storage phase) and FP64 x
    n32 = sum(r.stats["matvecs_mixed"] for r in results)
    t32 = sum(r.stats["seconds_matvec_mixed"] for r in results)
    n64 = matvecs - n32
    t64 = mv_seconds - t32
    roof = roofline(D, e_off, n64, t64, n32, t32, cached)
    # HBM bytes per matvec launch from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE passes over the same kernels (scaled by the copy-kernel
    # calibration), weighted by this run's FP64-x / FP32-x launch counts
    tj = Path(a.traffic_json) if a.traffic_json else ROOT / "profiles" / f"matvec_traffic_{a.config}.json"
    if tj.exists():
        tdat = json.loads(tj.read_text())
        per = {k: v["fetch_bytes"] + v["write_bytes"] for k, v in tdat.get("per_launch", {}).items()}
        kf, km = ("cached_fp64", "cached_mixed") if cached else ("fp64", "mixed")
        if kf in per and (km in per or n32 == 0):
            roof["traffic"] = (per[kf] * n64 + per.get(km, 0.0) * n32) / max(n64 + n32, 1)
            roof["traffic_source"] = str(tj.relative_to(ROOT))

    # comparison solves (outside the timed region): plain FP64 from a random
    # start, and the other storage mode from the same learned start
    compare = None
    if not a.no_compare and (a.precision != "fp64" or a.start != "random"):
        compare = {}
        runs = [("fp64_random", "fp64", False)]
        if model is not None:
            other = "mixed" if a.precision == "fp64" else "fp64"
            runs.append((f"{other}_learned", other, True))
        for tag, prec, use_model in runs:
            fresh_cache()
            tc = time.perf_counter()
            rc = solve(prec, use_model=use_model)
            device_synchronize()
            t_ref = time.perf_counter() - tc
            log(f"comparison {tag}: E0={rc.evals[0]:.15f} res={rc.rnorms[0]:.2e} "
                f"matvecs={rc.stats['matvecs']} t={t_ref:.3f}s")
            compare[f"{tag}_s"] = t_ref
            compare[f"{tag}_matvecs"] = rc.stats["matvecs"]
            compare[f"{tag}_E0"] = float(rc.evals[0])
            compare[f"{tag}_residual"] = float(rc.rnorms[0])
            compare[f"speedup_vs_{tag}"] = t_ref / per_step
        compare["E0_difference_vs_fp64_random"] = float(last.evals[0] - compare["fp64_random_E0"])

    # gather cache: built inside each solve; its build time apart (outside
    # the timed region): a matvec that builds it minus one that reads it
    cache_info = {"used": cached}
    if cached:
        spined.set_gather_cache(op, 0)
        spined.release_device_pool()
        xb = DeviceArray.from_numpy(np.ones(D))
        yb = DeviceArray(D, np.float64)
        spined.set_gather_cache(op, cache_mode)
        device_synchronize()
        t = time.perf_counter()
        spined.matvec_device(op, xb, yb, 1)
        device_synchronize()
        t_first = time.perf_counter() - t
        t = time.perf_counter()
        spined.matvec_device(op, xb, yb, 1)
        device_synchronize()
        t_cached = time.perf_counter() - t
        cache_info.update({"build_s": round(t_first - t_cached, 3),
                           "cached_fp64_matvec_ms": round(t_cached * 1e3, 2),
                           "entries_bytes_approx": e_off * 4})
        xb.free()
        yb.free()
        log(f"gather cache: {cache_info}")

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(n, basis, gens, lattice, a.cpu_sample_seconds, D,
                           matvecs / len(results))

    # the batched-solver workload beside the headline (its own JSON object):
    # the ring's device memory goes back first
    batch_res = None
    if not a.no_batch:
        spined.set_gather_cache(op, 0)
        del op, basis
        import gc
        gc.collect()
        spined.release_device_pool()
        batch_res = run_batch(a, CONFIGS["xxz512"][1], steps=1, warmup=0, with_cpu=not a.no_cpu)
        log(f"xxz512: {batch_res['value']:.1f} instances/s, launch {batch_res['result']['launch_s']:.2f} s")
    square_res = None
    if not a.no_square:
        import gc
        gc.collect()
        spined.release_device_pool()
        square_res = run_square6x6()
        log(f"square6x6: {square_res['result']}")

    e_exact = {40: -70.986091153332861, 24: -42.680058066148944}.get(n)
    out = {
        "metric": "s to E0 (residual<=1e-10), matvec HBM GB/s; mixed+learned vs plain FP64, 1 GPU",
        "value": per_step,
        "unit": "s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": per_step * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32-storage/f64-accumulate",
        "data": "synthetic (deterministic YAML-equivalent deck; random start vector seed 0x5EED)",
        "config": {"workload": desc, "dimension": D, "precision": a.precision,
                   "start": a.start, "tolerance": a.tol, "basis_build_s": round(t_build, 3),
                   # deck to E0 in one process: basis build, the training data
                   # (engine ground states of the small rings), training, one solve
                   # (the HDF5 writes of the representatives and eigenvector not counted)
                   "end_to_end_s": round(t_build + (train_info["data_s"] + train_info["train_s"]
                                                     if train_info else 0.0) + per_step, 3),
                   "autotune": tune,
                   "learned_start": train_info},
        "comparison": compare,
        "result": {"E0": float(last.evals[0]), "residual": float(last.rnorms[0]),
                   "E0_error_vs_bethe": (float(last.evals[0]) - e_exact) if e_exact else None,
                   "matvecs_per_solve": matvecs / len(results),
                   "matvecs_fp32_x": n32 / len(results),
                   "matvecs_fp64_x": n64 / len(results),
                   "start_vector_s": last.stats["seconds_start"],
                   "gather_cache": cache_info,
                   "matvec_ms": mv_avg * 1e3,
                   "matvec_rows_per_s": D / mv_avg if mv_avg > 0 else None,
                   "matvec_fraction_of_solve": mv_seconds / elapsed},
        "roofline": roof,
        "cpu_baseline": cpu,
        "batch_xxz512": batch_res,
        "block_square6x6": square_res,
    }
    print(json.dumps(out))
    return 0


def roofline(D, e_off, n64, t64, n32, t32, cached):
    """Roofline of the matvec by its algorithmic bytes per launch,
    B = D (8 + 4 + s_x + s_y) + E_off s_x: representative (8), stored
    stabiliser size (4), own x and y element, one x element per
    off-diagonal gather; s_y = 8 (y is FP64 in both modes), s_x = 8 or 4.
    The gather cache's 4-byte entries are index traffic, not algorithmic
    bytes: they are reported apart (index_bytes_per_launch), with the
    fraction they would add."""
    b64 = D * (8 + 4 + 8 + 8) + e_off * 8
    b32 = D * (8 + 4 + 4 + 8) + e_off * 4
    idx = e_off * 4 if cached else 0
    tot_b = b64 * n64 + b32 * n32
    tot_t = t64 + t32
    achieved = tot_b / tot_t if tot_t > 0 else 0.0
    per = {}
    for tag, b, nl, t in (("fp64_x", b64, n64, t64), ("fp32_x", b32, n32, t32)):
        if nl and t > 0:
            ms = t / nl * 1e3
            per[tag] = {"launches": nl, "ms_per_launch": ms, "algorithmic_bytes": b,
                        "achieved_GBps": b / (ms * 1e-3) / 1e9,
                        "frac": b / (ms * 1e-3) / PEAK_HBM,
                        "frac_with_index_bytes": (b + idx) / (ms * 1e-3) / PEAK_HBM}
    return {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": achieved / PEAK_HBM, "traffic": None,
            "kernel": "matvec_cached_kernel" if cached else "matvec_real_kernel",
            "algorithmic_bytes_per_launch": tot_b / max(n64 + n32, 1),
            "index_bytes_per_launch": idx,
            "offdiag_gathers_per_launch": e_off,
            "per_kernel": per}


# Krylov basis size of the batched solves (the library's default for
# spined_batch_eigsh; the byte model below uses the same value)
BATCH_BASIS = 16


def batch_roofline(sysb, res, m=BATCH_BASIS):
    """Roofline of the persistent batched launch by a model of its
    algorithmic bytes: per matvec of instance i, D_i (8 + 4 + 8 + 8) +
    E_i 8 (representative, stabiliser, own x, y, one x per off-diagonal term;
    E_i counted on the GPU per ring size), plus per Lanczos step the
    Gram-Schmidt and store passes, (2 j + 7) 8 D_i at the mean basis size
    j = (keep + m - 1) / 2, and per restart cycle (m + keep + 2) 8 D_i;
    divided by the launch time (HIP events around the one launch)."""
    from spined import internal as I
    keep = m // 2
    e_off = {}
    for x, op in zip(sysb.instances, sysb.ops):
        if x["n"] not in e_off:
            e_off[x["n"]] = I.count_offdiag(op)
    dims = sysb.dimensions
    tot = 0.0
    for x, D, mv, cy in zip(sysb.instances, dims, res["matvecs"], res["cycles"]):
        D = float(D)
        jbar = (keep + m - 1) / 2.0
        per_mv = D * 28 + e_off[x["n"]] * 8 + (2 * jbar + 7) * 8 * D
        tot += mv * per_mv + max(int(cy), 1) * (m + keep + 2) * 8 * D
    t = res["stats"]["seconds_cycles"]
    ach = tot / t if t > 0 else 0.0
    roof = {"bound": "hbm", "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": ach / PEAK_HBM, "traffic": None, "kernel": "batch_persistent_kernel",
            "algorithmic_bytes_per_launch": tot, "launch_s": t,
            "offdiag_terms_per_matvec_by_size": {str(k): v for k, v in sorted(e_off.items())}}
    # HBM bytes of the launch from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE passes over the same workload (copy-kernel calibration)
    tj = ROOT / "profiles" / "batch_traffic_xxz512.json"
    if tj.exists():
        tdat = json.loads(tj.read_text())
        if tdat.get("basis_size") == m:
            roof["traffic"] = tdat["fetch_bytes"] + tdat["write_bytes"]
            roof["traffic_source"] = str(tj.relative_to(ROOT))
    return roof


def run_square6x6():
    """Heisenberg 6x6 square lattice (72 bonds, w = 18, 36 translations x D4
    x spin inversion, trivial irrep, 15,804,956 states): lowest eigenvalue to
    residual 1e-10 by K = 8 block Lanczos with FP64 and with FP32 storage,
    beside the one-vector FP64 solve; each solve builds its own gather cache
    inside its time."""
    import spined
    from spined import lattice
    from spined.device import device_synchronize
    L = 6
    t = time.perf_counter()
    basis = spined.mk_basis(spined.mk_group([spined.mk_symmetry(p, s)
                                             for p, s in lattice.square_symmetries(L)]),
                            L * L, L * L // 2, 1)
    op = spined.mk_operator(basis, [spined.mk_interaction(lattice.HEISENBERG,
                                                          lattice.square_bonds(L))])
    spined.build_basis(basis)
    device_synchronize()
    t_build = time.perf_counter() - t
    res = {}
    for name, prec, blk, mb in (("block8_fp64", "fp64", 8, 48), ("block8_mixed", "mixed", 8, 48),
                                ("onevector_fp64", "fp64", 0, 0)):
        spined.set_gather_cache(op, 0)
        spined.set_gather_cache(op, -1)
        device_synchronize()
        t0 = time.perf_counter()
        r = spined.eigsh(op, k=1, tol=1e-10, precision=prec, max_block_size=blk,
                         max_basis_size=mb, want_vectors=False)
        device_synchronize()
        res[name] = {"s": round(time.perf_counter() - t0, 3), "E0": float(r.evals[0]),
                     "residual": float(r.rnorms[0]), "status": int(r.status),
                     "vector_matvecs": int(r.stats["matvecs"])}
    spined.set_gather_cache(op, 0)
    return {"metric": "s to E0 (residual<=1e-10), 6x6 square lattice, K=8 block Lanczos",
            "unit": "s", "higher_is_better": False,
            "config": {"workload": "Heisenberg 6x6 square lattice, 72 bonds, w=18, "
                                   "36 translations x D4 x spin inversion, trivial irrep",
                       "dimension": int(spined.get_number_states(basis)),
                       "basis_build_s": round(t_build, 3),
                       "E0_published": 4 * 36 * -0.678872},
            "result": res}


def run_batch(a, desc, steps=1, warmup=0, with_cpu=True):
    """Config 4: all 512 instances per step, one persistent launch per step."""
    from spined import batch
    from spined.device import device_synchronize
    t = time.perf_counter()
    inst = batch.xxz_instances(512)
    sysb = batch.BatchSystem(inst)
    device_synchronize()
    t_setup = time.perf_counter() - t
    dims = sysb.dimensions
    log(f"xxz512: {len(inst)} instances, sum of dimensions {int(dims.sum())}, setup {t_setup:.2f} s")
    for _ in range(warmup):
        batch.batch_eigsh(sysb.ops, tol=1e-9, max_basis_size=BATCH_BASIS)
    device_synchronize()
    t0 = time.perf_counter()
    res = None
    for _ in range(steps):
        res = batch.batch_eigsh(sysb.ops, tol=1e-9, max_basis_size=BATCH_BASIS)
        log(f"batch: {res['stats']}")
    device_synchronize()
    per = (time.perf_counter() - t0) / steps
    sec = res["seconds"]
    cpu = None
    if with_cpu:
        cpu = batch_cpu_baseline(inst, dims, a.cpu_sample_seconds, per, len(inst))
    return {
        "metric": "instances/s to residual<=1e-9 (batched uneven XXZ solves), 1 GPU",
        "value": len(inst) / per, "unit": "instances/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": per * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (numpy default_rng(20261202))",
        "config": {"workload": desc, "instances": len(inst), "sum_dimensions": int(dims.sum()),
                   "setup_s": round(t_setup, 2)},
        "result": {"converged": int(res["stats"]["converged"]),
                   "max_residual": float(np.max(res["rnorms"])),
                   "cycles_max": int(res["stats"]["cycles"]),
                   "launch_s": float(res["stats"]["seconds_cycles"]),
                   "p50_instance_s": float(np.percentile(sec, 50)),
                   "p99_instance_s": float(np.percentile(sec, 99)),
                   "max_matvecs_per_instance": int(res["stats"]["matvecs_per_instance_max"])},
        "roofline": batch_roofline(sysb, res),
        "cpu_baseline": cpu,
    }


def bench_batch(a, desc):
    steps = a.steps if a.steps is not None else 1
    warmup = a.warmup if a.warmup is not None else 0
    print(json.dumps(run_batch(a, desc, steps, warmup, not a.no_cpu)))
    return 0


def batch_cpu_baseline(inst, dims, budget_s, gpu_s, count):
    """Oracle Lanczos on all host cores, one instance at a time, over the
    smallest instances that fit the time budget; extrapolated by dimension."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    from spined import lattice
    cores = os.cpu_count() or 1
    order = np.argsort(dims)
    t_all = 0.0
    rows = 0
    done = 0
    for i in order:
        x = inst[i]
        n = x["n"]
        t = time.perf_counter()
        o = pyoracle.System(n, n // 2, 1, [])
        for k in range(n):
            o.add_interaction(np.asarray(lattice.xxz_matrix(x["delta"], x["J"][k]),
                                         dtype=np.complex128), [[k, (k + 1) % n]])
        o.build()
        o.lanczos(k=1, tol=1e-9, max_iter=300)
        t_all += time.perf_counter() - t
        rows += int(dims[i])
        done += 1
        if t_all > budget_s:
            break
    rate = rows / t_all
    est = float(np.sum(dims)) / rate
    return {"value": count / est, "unit": "instances/s (extrapolated)", "cores": cores,
            "kind": "port",
            "sample": f"{done} smallest instances solved one at a time by the oracle "
                      f"({rows} rows in {t_all:.1f} s); extrapolated by total dimension"}


def cpu_baseline(n, basis, gens, lattice, budget_s, D, matvecs_per_solve):
    """Oracle matvec on all host cores over a bounded row sample,
    extrapolated to one solve."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    import spined

    cores = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    reps = spined.basis_get_states(basis)
    stab = spined.basis_get_stabilizers(basis)
    o = pyoracle.System(n, n // 2, 1, gens)
    o.add_interaction(pyoracle.heisenberg_matrix(), lattice.ring_bonds(n))
    o.set_basis_with_stab(reps, stab)
    x = np.ones(D)
    # grow the sample until it costs about budget_s
    rows = min(D, 4096)
    start = D // 3
    elapsed = 0.0
    done_rows = 0
    while True:
        b = min(start, D - rows)
        t = time.perf_counter()
        o.matvec(x, rows=(b, b + rows))
        dt = time.perf_counter() - t
        elapsed += dt
        done_rows += rows
        if rows >= D or elapsed > budget_s / 3 or dt > budget_s / 6:
            break
        rows = min(D, rows * 4)
    rate = done_rows / elapsed
    t_matvec = D / rate
    return {"value": t_matvec * matvecs_per_solve, "unit": "s (extrapolated)", "cores": cores,
            "kind": "port",
            "sample": f"oracle matvec over {done_rows} rows of {D} ({elapsed:.1f} s, "
                      f"{rate:.3e} rows/s on {cores} threads); x {matvecs_per_solve:.0f} matvecs "
                      f"of the GPU solve, vector operations not counted"}


if __name__ == "__main__":
    sys.exit(main())
