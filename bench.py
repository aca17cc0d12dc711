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
PEAK_FP32_TFLOPS = 157.3  # dense FP32 (matrix) peak of MI355X
PEAK_HBM_GBS = 8000.0
NB = 64  # LU panel width (lu.hip)

# rocprofv3 kernel names of the kernel families nec_gpu_kernel_times reports
ROCPROF_NAME = {"lu_panel": "lu_panel_kernel", "zgemm": "zgemm_update_kernel", "lu_trsm": "lu_trsm_kernel",
                "lu_laswp": "lu_laswp_kernel", "fill_tile": "fill_tile_kernel", "fill_fields": "fill_fields_kernel",
                "fill_assemble": "fill_assemble_kernel", "pattern_partial": "pattern_partial_kernel",
                "trsv_lower": "trsv_blocked_kernel<true>", "trsv_upper": "trsv_blocked_kernel<false>"}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--nn", type=int, default=41, help="grid nodes per side (41 -> N=3,280)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-mixed", action="store_true", help="skip the N=16,020 mixed-vs-FP64 leg")
    p.add_argument("--no-batch", action="store_true", help="skip the batched Yagi-population leg")
    p.add_argument("--no-adjoint", action="store_true", help="skip the 256-design adjoint-gradient leg")
    p.add_argument("--mixed-nn", type=int, default=90, help="grid nodes per side of the mixed leg (90 -> N=16,020)")
    p.add_argument("--summarize", metavar="DIR", help="print a rocprofv3 --stats CSV summary and exit")
    p.add_argument("--summarize-pmc", nargs=2, metavar=("DIR", "COUNTER"),
                   help="print per-kernel means of one rocprofv3 --pmc counter and exit")
    p.add_argument("--copy-probe", type=int, metavar="MIB", help="run only a plain MIB-MiB device copy and exit")
    return p.parse_args()


def lu_work(n):
    """Algorithmic work of one factorisation, per kernel family.

    Complex multiply-add = 8 real flops, complex scaling by the pivot = 6;
    row exchanges move 16-byte entries (read + write)."""
    panel = gemm = trsm = 0.0
    laswp_bytes = 0.0
    for k0 in range(0, n, NB):
        nb = min(NB, n - k0)
        m = n - k0
        for c in range(nb):
            below = m - c - 1
            panel += 8.0 * below * (nb - c - 1) + 6.0 * below
        rest = n - k0 - nb
        laswp_bytes += 2.0 * 2.0 * 16.0 * nb * (n - nb)
        if rest > 0:
            trsm += 8.0 * rest * nb * (nb - 1) / 2.0
            gemm += 8.0 * rest * rest * nb
    return {"lu_panel": ("flop", panel), "zgemm": ("flop", gemm), "lu_trsm": ("flop", trsm),
            "lu_laswp": ("byte", laswp_bytes)}


def read_profile_value(name, key):
    """A "key=value" figure from a committed rocprofv3 summary under
    profiles/, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + "="):
            return float(line.split("=", 1)[1])
    return None


def read_traffic(kernel, workload_tag):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, or None."""
    return read_profile_value(f"{kernel}_{workload_tag}_traffic.txt", "traffic_bytes_per_launch")


def op_counts(workload_tag):
    """The oracle's operation tally of the fill and the pattern sum for a
    wire-grid workload (tests/golden/fill_counts.json, made by
    tests/golden/make_fill_counts.py), or None for another workload."""
    path = os.path.join(ROOT, "tests", "golden", "fill_counts.json")
    if not os.path.exists(path):
        return None
    return json.load(open(path)).get(workload_tag)


def fill_rates(counts, fill_kernel_ms, stage_ms, workload_tag=None, keys=("fill_tile",)):
    """Achieved FP64 rate of the fill from the tallied flops, and its HBM
    traffic (the one-pass fill kernel, or the field and assembly kernels
    when those ran) from the committed --pmc summary against the 16 N^2
    bytes of Z."""
    if counts is None:
        return None
    c = counts["fill"]
    tf = c["flops"] / (fill_kernel_ms * 1e-3) / 1e12
    traffic = None
    if workload_tag is not None:
        parts = [read_profile_value(f"fill_and_pattern_{workload_tag}_traffic.txt", f"{k}_traffic_bytes_per_launch")
                 for k in keys]
        traffic = None if None in parts else sum(parts)
    z_bytes = 16.0 * counts["n"] ** 2
    return {"hbm_traffic_bytes": traffic, "z_bytes": z_bytes,
            "traffic_over_z_bytes": None if traffic is None else traffic / z_bytes,
            "hbm_gbs": None if traffic is None else traffic / (fill_kernel_ms * 1e-3) / 1e9,
            "flops": c["flops"], "divs": c["divs"], "sqrts": c["sqrts"], "specials": c["specials"],
            "pairs": c["pairs"], "near_pairs": c["near_pairs"], "kernel_ms": fill_kernel_ms,
            "achieved_tflops": tf, "frac_fp64_vector_peak": tf / PEAK_FP64_TFLOPS,
            "specials_per_s": c["specials"] / (fill_kernel_ms * 1e-3),
            "stage_ms": stage_ms, "achieved_tflops_stage": c["flops"] / (stage_ms * 1e-3) / 1e12,
            "count": "oracle tally, tests/golden/fill_counts.json"}


def pattern_rates(counts, n, kernel_ms, workload_tag):
    """Achieved FP64 rate and HBM rates of the pattern's segment sums.

    Algorithmic bytes: the segment data read once (x, y, z, the three
    direction cosines, the length and the six current coefficients, 13
    doubles per segment) and the three complex sums written once per
    direction (48 bytes)."""
    if counts is None:
        return None
    c = counts["pattern"]
    n_dir = counts["pattern_directions"]
    alg_bytes = 13.0 * 8.0 * n + 48.0 * n_dir
    t = kernel_ms * 1e-3
    tf = c["flops"] / t / 1e12
    traffic = read_profile_value(f"fill_and_pattern_{workload_tag}_traffic.txt", "pattern_partial_traffic_bytes_per_launch")
    return {"flops": c["flops"], "divs": c["divs"], "specials": c["specials"], "terms": c["pairs"],
            "kernel_ms": kernel_ms, "achieved_tflops": tf, "frac_fp64_vector_peak": tf / PEAK_FP64_TFLOPS,
            "specials_per_s": c["specials"] / t,
            "algorithmic_bytes": alg_bytes, "achieved_gbs": alg_bytes / t / 1e9,
            "hbm_traffic_bytes": traffic, "hbm_gbs": None if traffic is None else traffic / t / 1e9,
            "frac_hbm_peak": None if traffic is None else traffic / t / 1e9 / PEAK_HBM_GBS,
            "count": "oracle tally, tests/golden/fill_counts.json"}


def short_name(name):
    """Kernel name without namespaces, return type and parameter list."""
    name = name.replace("(anonymous namespace)::", "").replace("necgpu::", "")
    if name.startswith("void "):
        name = name[5:]
    return name.split("(")[0]


def summarize(prof_dir):
    """Print the per-kernel statistics of a rocprofv3 --kernel-trace --stats -f csv run."""
    import csv
    import glob
    stats = sorted(glob.glob(os.path.join(prof_dir, "**", "*kernel_stats.csv"), recursive=True))
    if not stats:
        raise SystemExit(f"no *kernel_stats.csv under {prof_dir}")
    rows = list(csv.DictReader(open(stats[0])))
    print(f"{'kernel':44s} {'calls':>7s} {'total ms':>10s} {'avg us':>10s} {'%':>6s}")
    for r in rows:
        name = short_name(r["Name"])
        print(f"{name[:44]:44s} {int(r['Calls']):7d} {float(r['TotalDurationNs']) / 1e6:10.3f} "
              f"{float(r['AverageNs']) / 1e3:10.2f} {float(r['Percentage']):6.2f}")


def summarize_pmc(prof_dir, counter):
    """Per-kernel mean of one rocprofv3 --pmc counter (CSV output), in the counter's unit."""
    import csv
    import glob
    files = sorted(glob.glob(os.path.join(prof_dir, "**", "*counter_collection.csv"), recursive=True))
    if not files:
        raise SystemExit(f"no *counter_collection.csv under {prof_dir}")
    acc = {}
    for r in csv.DictReader(open(files[0])):
        if r.get("Counter_Name") != counter:
            continue
        name = short_name(r["Kernel_Name"])
        key = (name, r.get("Dispatch_Id", r.get("Correlation_Id")))
        acc[key] = acc.get(key, 0.0) + float(r["Counter_Value"])
    per = {}
    for (name, _), v in acc.items():
        per.setdefault(name, []).append(v)
    print(f"{'kernel':44s} {'dispatches':>10s} {counter + ' mean':>16s}")
    for name, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        print(f"{name[:44]:44s} {len(v):10d} {sum(v) / len(v):16.1f}")


def copy_probe(mib):
    """A plain device-to-device copy of known size (the --pmc scaling reference)."""
    import torch
    a = torch.ones(mib * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": a.numel() * 8, "copy_probe_bytes_written": b.numel() * 8}))


def solve_times(deck, precision, steps):
    """The on-device pipeline of the deck (fill, LU, solve, Z_in and, with an
    RP card, the pattern), timed step by step after one warm-up solve:
    (median ms per solve, Z_in, median stage ms, refinement, results) where
    results holds the currents and the total-gain pattern of the solve."""
    from necpp_amd import necpp as api
    import numpy as np
    ctx = api.nec_create()
    try:
        for w in deck["wires"]:
            api.nec_wire(ctx, *w)
        api.nec_geometry_complete(ctx, deck["gpflag"])
        if api.nec_gpu_set_precision(ctx, precision) != 0:
            raise RuntimeError(api.nec_error_message())
        api.nec_gpu_set_kernel_timing(ctx, 0)
        api.nec_fr_card(ctx, 0, 1, deck["freq_mhz"], 0.0)
        s, v = deck["source"], deck["voltage"]
        api.nec_ex_card(ctx, 0, s[0], s[1], 0, v.real, v.imag, 0, 0, 0, 0)
        rc = api.nec_rp_card(ctx, *deck["rp"]) if "rp" in deck else api.nec_xq_card(ctx, 0)
        if rc != 0:
            raise RuntimeError(api.nec_error_message())
        res = {"current": api.gpu_currents(ctx, 0)}
        if "rp" in deck:
            res["pg_tot"] = api.gpu_pattern_array(ctx, 0, "pg_tot", deck["rp"][1] * deck["rp"][2])
        ref = api.gpu_refinement(ctx) if precision == 1 else None
        out = np.zeros(8)
        per, stages = [], []
        for k in range(steps):
            if api.nec_gpu_benchmark_solve(ctx, 1 if k == 0 else 0, 1, out.ctypes.data) != 0:
                raise RuntimeError(api.nec_error_message())
            per.append(float(out[0]))
            stages.append([float(x) for x in out[1:5]])
        med = float(np.median(per))
        st = [float(x) for x in np.median(np.array(stages), axis=0)]
        res["step_ms"] = per
        return med, complex(out[5], out[6]), st, ref, res
    finally:
        api.nec_delete(ctx)


def pattern_errors(pg, ref, span_db=80.0):
    """Total power gain against the oracle's: largest |difference| in dB within
    span_db of the maximum, and the largest difference in linear gain
    relative to the maximum (nulls deeper than span_db are held to that)."""
    import numpy as np
    lin = lambda a: np.where(a > -999.0, 10.0 ** (a / 10.0), 0.0)
    live = ref > np.max(ref) - span_db
    return {"max_abs_db_within_%d_db" % int(span_db): float(np.max(np.abs(pg[live] - ref[live]))),
            "max_linear_rel_to_peak": float(np.max(np.abs(lin(pg) - lin(ref))) / np.max(lin(ref)))}


def mixed_leg(nn, steps=5):
    """FP64 vs mixed precision on the nn x nn grid with the 91 x 361 pattern,
    each timed as the median of `steps` solves; Z_in, currents and the
    pattern of both checked against the oracle fixture."""
    from necpp_amd import decks
    import numpy as np
    deck = decks.grid(nn)
    n = len(deck["wires"])
    ms64, z64, st64, _, r64 = solve_times(deck, 0, steps)
    ms32, z32, st32, (isteps, ires), r32 = solve_times(deck, 1, steps)
    fixture = os.path.join(ROOT, "tests", "golden", "c5_oracle.json")
    zref = None
    errors = None
    if n == 16020 and os.path.exists(fixture):
        zref = complex(*json.load(open(fixture))["zin"])
        golden = os.path.join(ROOT, "tests", "golden")
        cur = np.load(os.path.join(golden, "c5_oracle_current.npy"))
        pat = np.load(os.path.join(golden, "c5_oracle_pattern.npy"))
        errors = {}
        for tag, z, r in (("fp64", z64, r64), ("mixed", z32, r32)):
            errors[tag] = {"zin_rel": abs(z - zref) / abs(zref),
                           "currents_rel_normwise": float(np.linalg.norm(r["current"] - cur) / np.linalg.norm(cur)),
                           "pattern_total_gain": pattern_errors(r["pg_tot"], pat)}
    # the LU stage against the peaks: complex LU flops 8 n^3 / 3, FP64 factor
    # on the FP64 peak, the mixed path's FP32 factor on the FP32 peak
    lu_flops = 8.0 * float(n) ** 3 / 3.0
    lu64 = lu_flops / st64[1] / 1e9
    lu32 = lu_flops / st32[1] / 1e9
    return {"workload": f"wire grid {nn}x{nn} nodes, N={n}, fill+LU+solve+Z_in+91x361 pattern",
            "steps": steps, "statistic": "median",
            "fp64_ms_per_solve": ms64, "mixed_ms_per_solve": ms32, "speedup": ms64 / ms32,
            "fp64_step_ms": r64["step_ms"], "mixed_step_ms": r32["step_ms"],
            "fp64_lu_stage": {"achieved_tflops": lu64, "frac_fp64_peak": lu64 / PEAK_FP64_TFLOPS, "flops": lu_flops},
            "mixed_lu_stage": {"achieved_tflops": lu32, "frac_fp32_peak": lu32 / PEAK_FP32_TFLOPS, "flops": lu_flops},
            "fp64_stages_ms": dict(zip(("fill", "factor", "solve", "pattern"), st64)),
            "mixed_stages_ms": dict(zip(("fill", "factor", "solve", "pattern"), st32)),
            "refinement_steps": isteps, "refinement_rel_residual": ires,
            "zin_fp64": [z64.real, z64.imag], "zin_mixed": [z32.real, z32.imag],
            "zin_mixed_vs_fp64_rel": abs(z32 - z64) / abs(z64),
            "zin_mixed_vs_oracle_rel": None if zref is None else abs(z32 - zref) / abs(zref),
            "zin_fp64_vs_oracle_rel": None if zref is None else abs(z64 - zref) / abs(zref),
            "errors_vs_oracle": errors,
            "fp64_fill_stage": fill_rates(op_counts(f"grid{nn}"), st64[0], st64[0]),
            "fp64_pattern_stage": pattern_rates(op_counts(f"grid{nn}"), n, st64[3], f"grid{nn}")}


def batch_leg(repeats=3):
    """512 seeded Yagi designs x 64 frequencies with 361-point E-plane cuts, one batch."""
    from necpp_amd import decks
    from necpp_amd import necpp as api
    import numpy as np
    designs, freqs, theta, phi = decks.yagi_population()
    t = time.perf_counter()
    zin, gain, st = api.gpu_batch_solve(designs, freqs, theta, phi, repeats=repeats)
    wall_timed = time.perf_counter() - t
    # a production call: the same batch once (host preparation, upload, one
    # pass, results; the panel width found above is reused), its results
    # equal to the timed call's
    t = time.perf_counter()
    zin1, gain1, st1 = api.gpu_batch_solve(designs, freqs, theta, phi, repeats=0)
    wall = time.perf_counter() - t
    same = bool(np.array_equal(zin1, zin) and np.array_equal(gain1, gain))
    n = np.array([sum(w[1] for w in d["wires"]) for d in designs])
    lu_flops = float(np.sum(8.0 * n.astype(float) ** 3 / 3.0)) * len(freqs)
    return {"workload": "512 Yagi-Uda designs (seed %d, N = 63..252) x 64 frequencies 140-150 MHz, "
                        "fill+LU+solve+Z_in+361-point E-plane cut per solve" % decks.BATCH_SEED,
            "solves": int(st["solves"]), "ms_per_batch": st["graph_ms_per_replay"],
            "solves_per_s": st["solves"] / (st["graph_ms_per_replay"] * 1e-3),
            "stages_ms": {"fill": st["fill_ms"], "lu": st["lu_ms"], "solve_pattern": st["solve_pattern_ms"]},
            "lu_tflops": lu_flops / (st["lu_ms"] * 1e-3) / 1e12,
            "panel_width": int(st["panel_width"]), "graph_replays_equal_eager": st["graph_matches_eager"] == 1,
            "host_prep_ms": st1["host_ms"], "timed_call_wall_s": wall_timed,
            "end_to_end": {"solves_per_s": st["solves"] / wall, "call_wall_s": wall, "results_equal_timed_call": same,
                           "note": "one nec_gpu_batch_solve call without replays: host preparation, upload, "
                                   "one pass of the pipeline, results"}}


def adjoint_leg(count=256, single_sample=16):
    """J and dJ/dp (44 parameters) of 256 seeded 15-element Yagis (N = 465)
    by the adjoint method: all designs in one batched call
    (nec_gpu_batch_adjoint), plus a sample through the single-structure
    entry point (one design after another)."""
    from necpp_amd import adjoint as adj
    from necpp_amd import decks
    import numpy as np
    P = decks.c4_params(count=count)
    adj.batch_objective_and_gradient(P[:2])  # load the library's kernels once
    t = time.perf_counter()
    J, G, st = adj.batch_objective_and_gradient(P)
    wall = time.perf_counter() - t
    out = {"workload": "256 Yagi-Uda designs (15 elements x 31 segments, N = 465, 144 MHz, seed %d): "
                       "J = -G(90, 0) + 0.01 |Z_in - 50|^2 and its gradient in 44 parameters" % decks.ADJOINT_SEED,
           "designs": count, "wall_s": wall, "designs_per_s": count / wall, "batch_stats_ms": st}
    keys = ("solve_wall_ms", "adjoint_wall_ms", "device_ms", "host_ms", "fill_ms", "factor_ms", "solve_ms")
    acc = {k: 0.0 for k in keys}
    t = time.perf_counter()
    for p in P[:single_sample]:
        _, _, info = adj.objective_and_gradient(p)
        for k in keys:
            acc[k] += info[k]
    swall = time.perf_counter() - t
    out["single_path"] = {"designs": single_sample, "designs_per_s": single_sample / swall,
                          "mean_ms": {k: v / single_sample for k, v in acc.items()}}
    fx = os.path.join(ROOT, "tests", "golden", "c4_fd_oracle.npz")
    if os.path.exists(fx):
        z = np.load(fx)
        m = min(count, len(z["J"]))
        nf = np.linalg.norm(z["fd4"][:m], axis=1)
        err = np.linalg.norm(G[:m] - z["fd4"][:m], axis=1) / nf
        err_h = np.linalg.norm(G[:m] - z["fd"][:m], axis=1) / np.linalg.norm(z["fd"][:m], axis=1)
        out["vs_oracle_central_differences"] = {
            "step": "1e-6 |p|, fourth-order (4 fd(h) - fd(2h)) / 3, the fill's decisions held fixed",
            "designs": m, "within_1e-6": int((err <= 1e-6).sum()),
            "median_rel_err": float(np.median(err)), "max_rel_err": float(err.max()),
            "second_order_fd_h": {"within_1e-6": int((err_h <= 1e-6).sum()), "max_rel_err": float(err_h.max())},
            "J_max_rel_diff": float(np.max(np.abs(J[:m] - z["J"][:m]) / np.abs(z["J"][:m])))}
    return out


def cpu_baseline(deck, single_thread=True):
    """Time the oracle (CPU restatement) on one full solve of the same deck,
    on the host's cores (OpenMP) and on one thread (the reference as shipped
    is single-threaded)."""
    from oracle import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    t = time.perf_counter()
    r = po.run_deck(deck, nthreads=threads)
    wall = time.perf_counter() - t
    out = {"value": 1.0 / wall, "unit": "solves/s", "cores": threads, "kind": "port",
           "sample": f"1 full solve of the same N={len(deck['wires'])} grid (fill+LU+solve+91x361 pattern) "
                     f"on {threads} OpenMP threads; stage s: " +
                     ", ".join(f"{k}={v:.3f}" for k, v in r["timing"].items())}
    if single_thread:
        t = time.perf_counter()
        r1 = po.run_deck(deck, nthreads=1)
        wall1 = time.perf_counter() - t
        out["single_thread"] = {"value": 1.0 / wall1, "unit": "solves/s", "cores": 1,
                                "sample": "1 full solve of the same grid on one thread; stage s: " +
                                          ", ".join(f"{k}={v:.3f}" for k, v in r1["timing"].items())}
    return out, r


def main():
    a = parse()
    if a.summarize:
        summarize(a.summarize)
        return
    if a.summarize_pmc:
        summarize_pmc(*a.summarize_pmc)
        return
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N>1 is not supported "
                                   "(the solve is not sharded across GPUs)", "n_gpus": a.gpus}))
        sys.exit(2)
    from necpp_amd import decks
    from necpp_amd import necpp as api
    import numpy as np

    deck = decks.grid(a.nn)
    n = len(deck["wires"])
    ctx = api.nec_create()
    decks.apply_to_api(deck, api, ctx)  # one solve through the C API (loads HBM)
    z_api = complex(api.nec_impedance_real(ctx, 0), api.nec_impedance_imag(ctx, 0))
    # per-kernel durations (HIP events on every launch) from a few steps of
    # their own; the timed steps then run without those events
    kt_steps = max(1, min(a.steps, 5))
    out = np.zeros(8)
    if api.nec_gpu_benchmark_solve(ctx, 1, kt_steps, out.ctypes.data) != 0:
        raise RuntimeError(api.nec_error_message())
    ms_with_kernel_events = float(out[0])
    ktimes = api.gpu_kernel_times(ctx)
    api.nec_gpu_set_kernel_timing(ctx, 0)
    if api.nec_gpu_benchmark_solve(ctx, a.warmup, a.steps, out.ctypes.data) != 0:
        raise RuntimeError(api.nec_error_message())
    ms = float(out[0])
    fill_ms, lu_ms, solve_ms, pat_ms = (float(v) for v in out[1:5])
    z_bench = complex(out[5], out[6])
    api.nec_delete(ctx)

    lu_flops = 8.0 * n ** 3 / 3.0
    lu_tf = lu_flops / (lu_ms * 1e-3) / 1e12
    work = lu_work(n)
    kernels = {}
    for name, (launches, total) in ktimes.items():
        if launches == 0:
            continue
        e = {"launches_per_step": launches / kt_steps, "ms_per_step": total / kt_steps,
             "avg_us": 1000.0 * total / launches}
        if name in work:
            kind, amount = work[name]
            per_launch = amount / (launches / kt_steps)
            if kind == "flop":
                e["achieved_tflops"] = per_launch / (total / launches * 1e-3) / 1e12
            else:
                e["achieved_gbs"] = per_launch / (total / launches * 1e-3) / 1e9
        kernels[name] = e
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"]) if kernels else None
    dk = kernels.get(dom, {"avg_us": None, "launches_per_step": None})
    workload_tag = f"grid{a.nn}"
    if dom is None:
        roof = {"bound": None, "achieved": None, "peak": None, "unit": None, "frac": None}
    elif dom in work and work[dom][0] == "flop":
        achieved = dk["achieved_tflops"]
        roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS}
    elif dom in work:
        achieved = dk["achieved_gbs"]
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS}
    else:
        roof = {"bound": None, "achieved": None, "peak": None, "unit": None, "frac": None}
    counts = op_counts(workload_tag)
    fill_keys = tuple(k for k in ("fill_tile", "fill_fields", "fill_assemble") if k in kernels)
    fill_kernel_ms = sum(kernels[k]["ms_per_step"] for k in fill_keys)
    fill_stage = fill_rates(counts, fill_kernel_ms, fill_ms, workload_tag, fill_keys) if fill_kernel_ms > 0 else None
    pat_kernel_ms = kernels.get("pattern_partial", {}).get("ms_per_step")
    pattern_stage = pattern_rates(counts, n, pat_kernel_ms, workload_tag) if pat_kernel_ms else None
    hw_name = f"fill_and_pattern_{workload_tag}_fp64_counters.txt"
    for stage, keys in ((fill_stage, fill_keys), (pattern_stage, ("pattern_partial",))):
        if stage is None:
            continue
        hw = [read_profile_value(hw_name, f"{k}_fp64_flops_per_launch") for k in keys]
        if all(v is not None for v in hw):
            stage["rocprof_fp64_flops"] = sum(hw)
            stage["rocprof_fp64_tflops"] = sum(hw) / (stage["kernel_ms"] * 1e-3) / 1e12
            stage["rocprof_over_tally"] = sum(hw) / stage["flops"]
    roof["traffic"] = read_traffic(dom, workload_tag)
    roof["kernel"] = ROCPROF_NAME.get(dom, dom)
    roof["avg_launch_us"] = dk["avg_us"]
    roof["launches_per_step"] = dk["launches_per_step"]
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
        "roofline": roof,
        "lu_stage": {"achieved_tflops": lu_tf, "frac_fp64_peak": lu_tf / PEAK_FP64_TFLOPS,
                     "flops": lu_flops},
        "fill_stage": fill_stage,
        "pattern_stage": pattern_stage,
        "kernels": kernels,
        "kernel_timing": {"steps": kt_steps, "ms_per_step_with_kernel_events": ms_with_kernel_events},
        "mixed_vs_fp64": None,
    }
    if not a.no_mixed:
        result["mixed_vs_fp64"] = mixed_leg(a.mixed_nn)
    if not a.no_batch:
        result["batch"] = batch_leg()
    if not a.no_adjoint:
        result["adjoint"] = adjoint_leg()
    if not a.no_cpu:
        cb, r = cpu_baseline(deck)
        result["cpu_baseline"] = cb
        result["zin_vs_cpu_rel"] = abs(z_bench - r["zin"][0]) / abs(r["zin"][0])
    print(json.dumps(result))


if __name__ == "__main__":
    main()
