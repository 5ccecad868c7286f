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

Also reported:
  - the roofline of the solve kernel (its VALU rate: the algorithmic
    operation count over the solve kernel's HIP-event time on the library
    stream, against the FP64 vector peak, FP32 for the mixed form);
  - the error of the mixed form against the FP64 form on this run's own
    outputs, per output (up, down, net flux in W m-2, heating rate in K/day);
  - the ragged 10 000-column batch (levels 30..137, LW/SW 50/50), both forms;
  - the serving loop under Poisson arrivals at 50, 200, 1000 and 2000
    calls/s (schedules replayed by the library's own submitter thread), both
    forms, p50/p99 submit-to-done latency;
  - the reference's own CPU case (64 columns x 60 layers, LW clear) on the GPU;
  - CPU baselines: the reference Fortran (compiled from its own sources into
    oracle/_ref) on bounded samples of this workload and of the 64 x 60 LW
    clear case, on this host's cores (core counts stated).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision mixed|fp64]
                      [--no-autotune] [--serve-seconds S] [--no-cpu-baseline]
                      [--no-ragged] [--no-c1] [--tile R,B,S]

--tile fixes the solve tile of the --precision form (recompute 0/1, waves
per block, waves per SIMD) instead of the autotuned one; a profiled run uses
it to count the instructions of a given tile.
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
# calls): not measurable inside this script, so taken from the rocprofv3
# FETCH_SIZE and WRITE_SIZE runs of this script on the same build, each scaled
# by what the counter reads for a plain copy of known size; the summary, its
# commands and the scale factors are in TRAFFIC_SOURCE
TRAFFIC = {"mixed": 2.959e9, "fp64": 3.164e9}
TRAFFIC_SOURCE = "profiles/c2_solve_traffic.md"


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


def _host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(budget_s=15.0, case="c2"):
    """The reference Fortran (oracle/_ref, compiled from the reference's own
    sources) on this host: a bounded sample, one process per core (the
    reference's own way to use several cores: each caller process runs its
    column segments, docs/techguide/interface.tex:27-42), at most 16.

    case "c2": all-sky SW + LW McICA calls of 64 columns x 90 layers (this
    benchmark's workload in 64-column segments); case "c1": LW clear-sky calls
    of 64 columns x 60 layers (the reference's own CPU configuration)."""
    from socrates_amd import synth
    exe = os.path.join(ROOT, "oracle", "_ref", "soc_runes_case")
    data = os.path.join(ROOT, "socrates_amd", "data")
    if not os.path.exists(exe):
        return None
    total = _host_cores()
    cores = min(total, 16)
    ncol = 64
    srcs = ("sw", "lw") if case == "c2" else ("lw",)
    nlay = L if case == "c2" else 60
    with tempfile.TemporaryDirectory() as tmp:
        files = {}
        for src in srcs:
            nb = 6 if src == "sw" else 9
            c = synth.make_case(ncol, nlay, src, nb, 777 + nb, cloudy=case == "c2")
            files[src] = os.path.join(tmp, src + ".bin")
            c.write(files[src])
        mc = os.path.join(data, "mcica_synth") if case == "c2" else "-"

        def run(src, reps, n):
            # the reference keeps its per-call work arrays on the stack: a
            # 64-column call needs more than the usual 8 MiB
            procs = [subprocess.Popen([exe, os.path.join(data, "sp_sw_synth"),
                                       os.path.join(data, "sp_lw_synth"), mc, files[src],
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
        t1 = sum(run(src, 1, 1) for src in srcs)
        reps = max(1, int(budget_s / max(t1, 1e-3)))
        t = sum(run(src, reps, cores) for src in srcs)
    solves = cores * reps * ncol * sum(G[src] for src in srcs) * nlay / 90.0
    what = ("all-sky SW+LW McICA, 90 layers" if case == "c2" else "LW clear sky, 60 layers")
    return {"value": solves / t, "unit": "column.g-point solves/s at 90 levels", "cores": cores,
            "host_cores": total, "kind": "reference",
            "sample": "%d processes x %d calls x %d columns, %s" % (cores, reps, ncol, what)}


def serving_lines(seconds, rates=(50.0, 200.0, 1000.0, 2000.0), precisions=("mixed", "fp64"),
                  seed=20261208):
    """The serving loop (socgpu_serve_*) under Poisson arrivals: each call
    carries 1..4096 columns (log-uniform), 40..137 layers, LW or SW, all-sky
    McICA, inputs resident in HBM; the arrival schedule is replayed by the
    library's own submitter thread (socgpu_serve_replay), so the offered rate
    does not depend on this interpreter."""
    import torch
    from socrates_amd import runes as R
    from socrates_amd import synth
    from socrates_amd.serve import Server
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    pool, keep = [], []
    for k in range(24):
        src = "lw" if k % 2 == 0 else "sw"
        n = max(1, int(np.exp(rng.uniform(0.0, np.log(4096.0)))))
        c = synth.make_case(n, int(rng.integers(40, 138)), src, 9 if src == "lw" else 6, seed + k,
                            cloudy=True, first_column_id=100000 * k)
        arrs = {k2: torch.from_numpy(np.ascontiguousarray(getattr(c, k2))).to(dev)
                for k2 in ("p_layer", "t_layer", "mass", "density", "layer_heat_capacity", "h2o", "o3",
                           "p_level", "t_level", "t_ground", "solar_irrad", "albedo_diff", "albedo_dir",
                           "cloud_frac", "liq_mmr", "liq_dim", "rand_seed")}
        arrs["cos_zenith_angle"] = torch.from_numpy(np.ascontiguousarray(c.cos_zen)).to(dev)
        keep.append(arrs)
        pool.append((src, c, R.case_call(c, src, device_arrays=arrs)))
    out = []
    for prec in precisions:
        for rate in rates:
            n = int(max(20, min(4000, rate * seconds)))
            arrival = np.cumsum(rng.exponential(1.0 / rate, n))
            picks = rng.integers(0, len(pool), n).astype(np.int32)
            srv = Server(precision=prec, hbm_budget_gb=96.0, n_slots=2, max_batch_requests=512)
            try:
                srv.wait(srv.submit(*pool[0][2]))      # warm the slots' buffers
                srv.reset_stats()
                wall = srv.replay([p[2] for p in pool], arrival, picks)
                st = srv.stats()
            finally:
                srv.close()
            solves = sum(pool[j][1].n_profile * G[pool[j][0]] * pool[j][1].n_layer / 90.0 for j in picks)
            out.append({"precision": prec, "rate_req_s": rate, "requests": n,
                        "offered_s": float(arrival[-1]), "wall_s": wall,
                        "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "max_ms": st["max_ms"],
                        "mean_batch": st["mean_batch"], "batches": st["batches"],
                        "solves_per_s_l90": solves / wall})
    return out


def c1_gpu_line(torch, dev, calls_per_rep=200):
    """The reference's own CPU configuration on the GPU: LW clear sky,
    64 columns x 60 layers, inputs and outputs in HBM, one call at a time."""
    from socrates_amd import runes as R
    from socrates_amd import synth
    c = synth.make_case(64, 60, "lw", 9, 20261202 + 1001, cloudy=False)
    arrs = {k: torch.from_numpy(np.ascontiguousarray(getattr(c, k))).to(dev)
            for k in ("p_layer", "t_layer", "mass", "density", "layer_heat_capacity", "h2o", "o3",
                      "p_level", "t_level", "t_ground", "solar_irrad", "albedo_diff", "albedo_dir")}
    arrs["cos_zenith_angle"] = torch.from_numpy(np.ascontiguousarray(c.cos_zen)).to(dev)
    res = {}
    for prec in ("fp64", "mixed"):
        for _ in range(5):
            R.runes_case(c, "lw", precision=prec, device_arrays=arrs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(calls_per_rep):
            R.runes_case(c, "lw", precision=prec, device_arrays=arrs)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / calls_per_rep
        res[prec] = {"ms_per_call": dt * 1e3, "solves_per_s_l90": 64 * G["lw"] * (60 / 90.0) / dt}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-ragged", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--tile", default=None,
                    help="recompute,waves_per_block,waves_per_simd for the --precision form")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain 512 MiB device copy, the known-size kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the path shards by columns with no collective, "
              "so only --gpus 1 is run here", file=sys.stderr)
        sys.exit(2)
    import torch
    from socrates_amd import errstats
    from socrates_amd import runes as R
    dev = torch.device("cuda:0")
    R.init(0)
    R.register_default_spectra()
    cases = make_inputs(torch, dev)
    if a.copy_calibration:
        src_buf = torch.ones(64 << 20, dtype=torch.float64, device=dev)
        dst_buf = torch.empty_like(src_buf)
        dst_buf.copy_(src_buf)
        torch.cuda.synchronize()
        del src_buf, dst_buf
    outs = {}
    for prec in ("fp64", "mixed"):
        outs[prec] = {}
        for src, (c, arrs) in cases.items():
            L1 = L + 1
            outs[prec][src] = dict(
                flux_up=torch.empty((L1, P), dtype=torch.float64, device=dev),
                flux_down=torch.empty((L1, P), dtype=torch.float64, device=dev),
                flux_direct=torch.empty((L1, P), dtype=torch.float64, device=dev),
                heating_rate=torch.empty((L, P), dtype=torch.float64, device=dev))

    precs = [a.precision] + (["fp64"] if a.precision != "fp64" else [])
    tiles = {}
    results = {}
    for prec in precs:
        # one step = one SW and one LW radiation call, packed into one batch;
        # the calls are marshalled once, as a host model sets up its calls
        calls = [R.case_call(cases[src][0], src, device_arrays=cases[src][1], out=outs[prec][src])
                 for src in ("sw", "lw")]
        reqs = R.Requests(calls)
        if a.tile and prec == a.precision:
            R.set_tile(prec, *[int(v) for v in a.tile.split(",")])
            tiles[prec] = R.get_tile(prec)
        elif not a.no_autotune:
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
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK[a.precision],
                     "unit": "TFLOP/s", "frac": achieved / PEAK[a.precision],
                     "traffic": TRAFFIC.get(a.precision), "traffic_source": TRAFFIC_SOURCE,
                     "engine": "VALU: the path has no matrix-shaped work; the compute roof "
                               "is the published FP64 vector rate (FP32 for the mixed form)",
                     "kernel": "b_solve_kernel", "kernel_ms_per_step": kt * 1e3,
                     "hbm_achieved_GBps": hbm, "hbm_frac": hbm / HBM_PEAK},
    }
    if a.precision != "fp64":
        line["fp64_value"] = solves_per_step / (results["fp64"]["dt"] / a.steps)
        line["mixed_over_fp64"] = value / line["fp64_value"]
        line["fp64_solve_kernel_ms_per_step"] = results["fp64"]["solve_ms"]
        line["fp64_tile"] = tiles["fp64"]
        # the mixed form against the FP64 form on this run's own outputs
        torch.cuda.synchronize()
        err = {}
        for src in ("sw", "lw"):
            st = errstats.stats(outs["mixed"][src], outs["fp64"][src])
            err[src] = {k: {"max": v["max"], "p99": v["p99"]} for k, v in st.items()
                        if k in ("flux_up", "flux_down", "flux_net", "heating_rate_kday")}
        line["mixed_vs_fp64_error"] = dict(err, units="flux W m-2, heating rate K/day",
                                           bounds={"flux": 0.05, "heating_rate_kday": 0.01})
    if not a.no_ragged:
        from socrates_amd import benchmarks
        rg = benchmarks.ragged(steps=5, warmup=2, quiet=True)
        line["ragged"] = {"workload": "10000 columns, levels U{30..137}, LW/SW 50/50, all-sky McICA, "
                                      "one graph-replayed batch of %d calls" % rg["calls_per_batch"],
                          "fp64_solves_per_s_l90": rg["fp64"]["solves_per_s_l90"],
                          "mixed_solves_per_s_l90": rg["mixed"]["solves_per_s_l90"],
                          "mixed_over_fp64": rg["mixed_over_fp64"]}
    if not a.no_c1:
        line["c1_gpu"] = dict(c1_gpu_line(torch, dev),
                              workload="64 columns x 60 layers, LW clear sky, one call at a time")
    if a.serve_seconds > 0:
        line["serving"] = serving_lines(a.serve_seconds)
    if not a.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline()
        except Exception as exc:  # reported, not fatal: the GPU figure stands on its own
            line["cpu_baseline"] = {"value": None, "error": str(exc)[:300]}
        if not a.no_c1:
            try:
                line["cpu_baseline_c1"] = cpu_baseline(budget_s=5.0, case="c1")
            except Exception as exc:
                line["cpu_baseline_c1"] = {"value": None, "error": str(exc)[:300]}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
