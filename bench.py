#!/usr/bin/env python3
"""Benchmark: solves/s of the N = 3,280 wire grid (FP64) on one MI355X.

One step = one complete solve of the 41 x 41-node, 0.05 m wire grid at
299.792458 MHz (N = 3,280 segments): impedance fill, LU, forward/back
solve, current coefficients, Z_in, and the 91 x 361 far-field pattern,
all on the device with the structure already resident in HBM
(nec_gpu_benchmark_solve).  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s (fill+LU+RP) at N=3,280 FP64; mixed-vs-FP64 speedup @1e-10 rel Zin"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--nn", type=int, default=41, help="grid nodes per side (41 -> N=3,280)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    return p.parse_args()


def cpu_baseline(deck):
    """Time the oracle (CPU restatement) on one full solve of the same deck."""
    from oracle import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    t = time.perf_counter()
    r = po.run_deck(deck, nthreads=threads)
    wall = time.perf_counter() - t
    return {"value": 1.0 / wall, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"1 full solve of the same N={len(deck['wires'])} grid (fill+LU+solve+91x361 pattern) "
                      f"on {threads} OpenMP threads; stage s: " +
                      ", ".join(f"{k}={v:.3f}" for k, v in r["timing"].items())}, r


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N>1 is not supported (see DESIGN.md "
                                   "for how the solve would shard)", "n_gpus": a.gpus}))
        sys.exit(2)
    from necpp_amd import decks
    from necpp_amd import necpp as api
    import numpy as np

    deck = decks.grid(a.nn)
    n = len(deck["wires"])
    ctx = api.nec_create()
    decks.apply_to_api(deck, api, ctx)  # one solve through the C API (loads HBM)
    z_api = complex(api.nec_impedance_real(ctx, 0), api.nec_impedance_imag(ctx, 0))
    out = np.zeros(8)
    if api.nec_gpu_benchmark_solve(ctx, a.warmup, a.steps, out.ctypes.data) != 0:
        raise RuntimeError(api.nec_error_message())
    ms = float(out[0])
    fill_ms, lu_ms, solve_ms, pat_ms = (float(v) for v in out[1:5])
    z_bench = complex(out[5], out[6])
    api.nec_delete(ctx)

    lu_flops = 8.0 * n ** 3 / 3.0
    lu_tf = lu_flops / (lu_ms * 1e-3) / 1e12
    result = {
        "metric": METRIC,
        "value": 1000.0 / ms,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"wire grid {a.nn}x{a.nn} nodes, N={n}, 299.792458 MHz, FP64 fill+LU+solve+Z_in+"
                               "91x361 pattern", "n_segments": n, "parallelism": "single GPU"},
        "stages_ms": {"fill": fill_ms, "factor": lu_ms, "solve": solve_ms, "pattern": pat_ms},
        "zin": [z_bench.real, z_bench.imag],
        "zin_api_rel_diff": abs(z_bench - z_api) / abs(z_api),
        "roofline": {"bound": "mfma", "achieved": lu_tf, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": lu_tf / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "LU stage (8N^3/3 flops per factorisation)"},
    }
    if not a.no_cpu:
        cb, r = cpu_baseline(deck)
        result["cpu_baseline"] = cb
        result["zin_vs_cpu_rel"] = abs(z_bench - r["zin"][0]) / abs(r["zin"][0])
    print(json.dumps(result))


if __name__ == "__main__":
    main()
