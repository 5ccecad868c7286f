#!/usr/bin/env python3
"""Benchmark of the two-stream flux path on one MI355X.

Workload (one step = one radiation time step of a host model): an all-sky
SW call (6 bands, 40 g-points) and an all-sky LW call (9 bands, 50 g-points)
on 4096 columns x 90 layers, McICA with 64 generated sub-columns and optimal
sampling, exponential-random overlap; synthetic columns and spectral files
(socrates_amd/synth.py, socrates_amd/data). Inputs and outputs stay in HBM.

Metric: column.g-point flux solves per second at 90 levels (one solve = one
(column, major k-term) flux profile including its sampled sub-columns).

The step runs as one packed batch (socgpu_batch_run: the SW and LW calls
in one launch sequence, replayed from a captured HIP graph) with the solve
tile picked by the autotuner (socgpu_autotune). The default form is the
mixed FP32/FP64 one; the FP64 form is timed beside it (fp64_value,
mixed_over_fp64).

Also reported: the roofline of the solve kernel (FP64/FP32 VALU rate from
the algorithmic operation count over the solve kernel's HIP-event time on
the library stream), a short serving run (Poisson arrivals, p50/p99
submit-to-done latency) and a CPU baseline: the reference Fortran (compiled
from its own sources into oracle/_ref) on a bounded sample, on this host's
cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision mixed|fp64]
                      [--no-autotune] [--serve-seconds S] [--no-cpu-baseline]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

P, L = 4096, 90
G = {"sw": 40, "lw": 50}
# sampled McICA sub-columns per column (optimal sampling of the McICA data
# file: subcol_need_optimal = 52, one per (band, k-term) item)
NSUB = {"sw": 52, "lw": 52}
# algorithmic operations per layer of one two-stream profile solve (a clear
# k-term or one sampled sub-column), counted from the reference formulas
# (exp, sqrt and div count as one op)
OPS_PER_LAYER = {"sw": 94.0, "lw": 79.0}
# compulsory HBM bytes per (column, g-point) solve: FP64 column inputs read
# once and outputs written once, divided by the g-points of the file
BYTES_PER_SOLVE = {"sw": 235.0, "lw": 188.0}
PEAK = {"fp64": 78.6, "mixed": 157.3}      # TFLOP/s vector
HBM_PEAK = 8000.0                          # GB/s
# HBM bytes per step moved by the solve kernel (one launch covers both
# calls), from rocprofv3 FETCH_SIZE (x2, the factor a plain 512 MiB copy
# reads) and WRITE_SIZE (x1) runs of this script; profiles/c2_solve_traffic.md
TRAFFIC = {"mixed": 2.134e9, "fp64": 4.260e9}


def make_inputs(torch, dev):
    from socrates_amd import synth
    cases = {}
    for src, nb, seed in (("sw", 6, 20261202 + 2), ("lw", 9, 20261202 + 1)):
        c = synth.make_case(P, L, src, nb, seed, cloudy=True)
        arrs = {}
        for k in ("p_layer", "t_layer", "mass", "density", "layer_heat_capacity", "h2o", "o3",
                  "p_level", "t_level", "t_ground", "solar_irrad", "albedo_diff", "albedo_dir",
                  "cloud_frac", "liq_mmr", "liq_dim"):
            arrs[k] = torch.from_numpy(np.ascontiguousarray(getattr(c, k))).to(dev)
        arrs["cos_zenith_angle"] = torch.from_numpy(np.ascontiguousarray(c.cos_zen)).to(dev)
        arrs["rand_seed"] = torch.from_numpy(np.ascontiguousarray(c.rand_seed)).to(dev)
        cases[src] = (c, arrs)
    return cases


def _raise_stack_limit():
    import resource
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    want = resource.RLIM_INFINITY if hard == resource.RLIM_INFINITY else hard
    resource.setrlimit(resource.RLIMIT_STACK, (want, hard))


def cpu_baseline(budget_s=15.0):
    """Reference Fortran on this host: bounded sample, one process per core."""
    from socrates_amd import synth
    exe = os.path.join(ROOT, "oracle", "_ref", "soc_runes_case")
    data = os.path.join(ROOT, "socrates_amd", "data")
    kind = "reference"
    if not os.path.exists(exe):
        return None
    cores = min(os.cpu_count() or 1, 16)
    ncol = 64
    with tempfile.TemporaryDirectory() as tmp:
        files = {}
        for src, nb in (("sw", 6), ("lw", 9)):
            c = synth.make_case(ncol, L, src, nb, 777 + nb, cloudy=True)
            files[src] = os.path.join(tmp, src + ".bin")
            c.write(files[src])

        def run(src, reps, n):
            # the reference keeps its per-call work arrays on the stack: a
            # 64-column call needs more than the usual 8 MiB
            procs = [subprocess.Popen([exe, os.path.join(data, "sp_sw_synth"),
                                       os.path.join(data, "sp_lw_synth"),
                                       os.path.join(data, "mcica_synth"), files[src],
                                       os.path.join(tmp, "%s%d.out" % (src, i)), str(reps)],
                                      stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                      text=True, preexec_fn=_raise_stack_limit)
                     for i in range(n)]
            secs = []
            for p in procs:
                out, _ = p.communicate()
                if p.returncode != 0 or "seconds=" not in out:
                    raise RuntimeError("reference harness failed (rc %d): %s"
                                       % (p.returncode, out[-300:]))
                secs.append(float(out.split("seconds=")[1]))
            return max(secs)
        # calibrate one call per source, then size the sample to the budget
        t1 = run("sw", 1, 1) + run("lw", 1, 1)
        reps = max(1, int(budget_s / max(t1, 1e-3)))
        t = run("sw", reps, cores) + run("lw", reps, cores)
    solves = cores * reps * ncol * (G["sw"] + G["lw"])
    return {"value": solves / t, "unit": "column.g-point solves/s", "cores": cores,
            "kind": kind,
            "sample": "%d processes x %d steps x %d columns x 90 layers, all-sky SW+LW McICA"
                      % (cores, reps, ncol)}


def serving_line(seconds, rate=200.0, precision="mixed", seed=20261208):
    """A short serving run (socrates_amd.serve): Poisson arrivals of calls
    with 1..4096 columns (log-uniform), 40..137 layers, LW or SW, all-sky."""
    import queue
    import threading
    from socrates_amd import runes as R
    from socrates_amd import synth
    from socrates_amd.serve import Server
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(16):
        src = "lw" if k % 2 == 0 else "sw"
        n = int(np.exp(rng.uniform(0.0, np.log(4096.0))))
        pool.append((src, synth.make_case(max(n, 1), int(rng.integers(40, 138)), src, 9 if src == "lw" else 6,
                                          seed + k, cloudy=True, first_column_id=100000 * k)))
    srv = Server(precision=precision, hbm_budget_gb=64.0, n_slots=2)
    srv.wait(srv.submit(*R.case_call(pool[0][1], pool[0][0])))
    srv.reset_stats()
    n = int(rate * seconds)
    gaps = rng.exponential(1.0 / rate, n)
    picks = rng.integers(0, len(pool), n)
    q = queue.Queue()
    th = threading.Thread(target=lambda: [srv.wait(t) for t in iter(q.get, None)], daemon=True)
    th.start()
    t0 = time.perf_counter()
    t_next = t0
    solves = 0.0
    for i in range(n):
        t_next += gaps[i]
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        src, c = pool[picks[i]]
        q.put(srv.submit(*R.case_call(c, src)))
        solves += c.n_profile * G[src] * c.n_layer / 90.0
    q.put(None)
    th.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return {"rate_req_s": rate, "requests": n, "precision": precision, "p50_ms": st["p50_ms"],
            "p99_ms": st["p99_ms"], "mean_batch": st["mean_batch"],
            "solves_per_s_l90": solves / wall}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the path shards by columns with no collective, "
              "so only --gpus 1 is run here", file=sys.stderr)
        sys.exit(2)
    import torch
    from socrates_amd import runes as R
    dev = torch.device("cuda:0")
    R.init(0)
    R.register_default_spectra()
    cases = make_inputs(torch, dev)
    outs = {}
    for src, (c, arrs) in cases.items():
        L1 = L + 1
        outs[src] = dict(flux_up=torch.empty((L1, P), dtype=torch.float64, device=dev),
                         flux_down=torch.empty((L1, P), dtype=torch.float64, device=dev),
                         flux_direct=torch.empty((L1, P), dtype=torch.float64, device=dev),
                         heating_rate=torch.empty((L, P), dtype=torch.float64, device=dev))
    # one step = one SW and one LW radiation call, packed into one batch
    calls = [R.case_call(cases[src][0], src, device_arrays=cases[src][1], out=outs[src])
             for src in ("sw", "lw")]
    reqs = R.Requests(calls)   # marshalled once, as a host model sets up its calls

    precs = [a.precision] + (["fp64"] if a.precision != "fp64" else [])
    tiles = {}
    results = {}
    for prec in precs:
        if not a.no_autotune:
            tiles[prec] = R.autotune(calls, prec)
        else:
            tiles[prec] = R.get_tile(prec)
        # solve-kernel time (HIP events on the library stream) from the
        # stream form of the same batch
        ks = []
        for _ in range(3):
            R.runes_batch(calls, precision=prec)
            ks.append(R.last_timing()[1])
        solve_ms = sorted(ks)[1]
        b = R.Batch(prec)
        for _ in range(a.warmup):
            b.run(reqs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev_ms = 0.0
        for _ in range(a.steps):
            b.run(reqs)
            dev_ms += b.stats()["last_ms"]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = b.stats()
        b.close()
        results[prec] = dict(dt=dt, solve_ms=solve_ms, graph=st, dev_ms=dev_ms / a.steps)
    main_r = results[a.precision]
    ms_per_step = main_r["dt"] / a.steps * 1e3
    solves_per_step = P * (G["sw"] + G["lw"])
    value = solves_per_step / (main_r["dt"] / a.steps)
    # roofline of the solve kernel (SW and LW calls of one step together):
    # every clear k-term and sampled sub-column profile of the reference's
    # mcica_sample (columns entirely clear or overcast skip one of the two,
    # here and in the reference, so this slightly overstates the work)
    flops = sum(P * (G[s] + NSUB[s]) * L * OPS_PER_LAYER[s] for s in G)
    bytes_alg = sum(P * G[s] * BYTES_PER_SOLVE[s] for s in G)
    kt = main_r["solve_ms"] * 1e-3
    achieved = flops / kt / 1e12
    hbm = bytes_alg / kt / 1e9
    line = {
        "metric": "column.g-point flux solves/s at 90 levels (all-sky SW+LW McICA)",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_per_step, "device_ms_per_step": main_r["dev_ms"],
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic columns and spectral files (seeded)",
        "config": {"workload": "4096 columns x 90 layers, all-sky SW (6 bands, 40 g) + LW "
                               "(9 bands, 50 g), McICA 64 sub-columns, exp-random overlap; "
                               "one step = the SW and LW calls as one graph-replayed batch",
                   "precision": a.precision, "tile": tiles[a.precision]},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK[a.precision],
                     "unit": "TFLOP/s", "frac": achieved / PEAK[a.precision],
                     "traffic": TRAFFIC.get(a.precision),
                     "engine": "VALU: the path has no matrix-shaped work; the compute roof "
                               "is the published FP64 (FP32 for the mixed form) rate",
                     "kernel": "b_solve_kernel", "kernel_ms_per_step": kt * 1e3,
                     "hbm_achieved_GBps": hbm, "hbm_frac": hbm / HBM_PEAK},
    }
    if a.precision != "fp64":
        line["fp64_value"] = solves_per_step / (results["fp64"]["dt"] / a.steps)
        line["mixed_over_fp64"] = value / line["fp64_value"]
        line["fp64_solve_kernel_ms_per_step"] = results["fp64"]["solve_ms"]
        line["fp64_tile"] = tiles["fp64"]
    if a.serve_seconds > 0:
        line["serving"] = serving_line(a.serve_seconds, precision=a.precision)
    if not a.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline()
        except Exception as exc:  # reported, not fatal: the GPU figure stands on its own
            line["cpu_baseline"] = {"value": None, "error": str(exc)[:300]}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
