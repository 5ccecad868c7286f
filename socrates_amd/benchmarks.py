"""Benchmarks beyond the headline one (bench.py at the repository root):

  ragged  - 10 000 columns with layer counts uniform in 30..137, half LW
            (9 bands), half SW (6 bands), all-sky McICA, grouped into one
            radiation call per (source, layer count) and run as one packed,
            graph-captured batch per step (socgpu_batch_run); FP64 and mixed.
  serve   - the serving loop under Poisson arrivals: each call carries 1..4096
            columns (log-uniform) with 40..137 layers, LW or SW, all-sky McICA;
            p50/p99 submit-to-done latency and throughput per arrival rate.
  triple  - stratiform + convective cloud (triple-column regions) on 4096
            columns x 90 layers, SW and LW in one graph-captured batch.
  ga9     - the production GA9 spectral files and McICA data on 4096 x 90
            all-sky columns, SW and LW in one graph-captured batch.

Inputs are synthetic (socrates_amd/synth.py, seeded). Throughput is in
column.g-point solves/s, and also normalised to 90 layers (x L/90).

    python -m socrates_amd.benchmarks ragged [--steps 10]
    python -m socrates_amd.benchmarks serve [--rates 50,200,1000,2000] [--seconds 10] [--inputs host|device]
    python -m socrates_amd.benchmarks triple [--steps 10]
    python -m socrates_amd.benchmarks ga9 [--steps 10]
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from socrates_amd import runes as R  # noqa: E402
from socrates_amd import synth  # noqa: E402

NBAND = {"lw": 9, "sw": 6}
NG = {"lw": 50, "sw": 40}
FIELDS = ("p_layer", "t_layer", "mass", "density", "layer_heat_capacity", "h2o", "o3", "p_level",
          "t_level", "t_ground", "solar_irrad", "albedo_diff", "albedo_dir", "cloud_frac", "liq_mmr",
          "liq_dim", "rand_seed")


def _device_case(torch, case, dev):
    arrs = {k: torch.from_numpy(np.ascontiguousarray(getattr(case, k))).to(dev) for k in FIELDS}
    arrs["cos_zenith_angle"] = torch.from_numpy(np.ascontiguousarray(case.cos_zen)).to(dev)
    L, P = case.n_layer, case.n_profile
    out = {k: torch.empty(((L + 1) if k != "heating_rate" else L, P), dtype=torch.float64, device=dev)
           for k in ("flux_up", "flux_down", "flux_direct", "heating_rate")}
    return arrs, out


def ragged(steps=10, warmup=2, n_columns=10000, seed=20261206, quiet=False):
    import torch
    dev = torch.device("cuda:0")
    R.init(0)
    R.register_default_spectra()
    rng = np.random.default_rng(seed)
    levels = rng.integers(30, 138, n_columns)
    source = np.where(np.arange(n_columns) < n_columns // 2, "lw", "sw")
    calls, keep = [], []
    solves = 0.0
    for src in ("lw", "sw"):
        for L in np.unique(levels[source == src]):
            P = int(((levels == L) & (source == src)).sum())
            case = synth.make_case(P, int(L), src, NBAND[src], seed=seed + 1000 * int(L) + (src == "sw"),
                                   cloudy=True, first_column_id=int(L) * 100000)
            arrs, out = _device_case(torch, case, dev)
            keep.append((arrs, out))
            calls.append(R.case_call(case, src, device_arrays=arrs, out=out))
            solves += P * NG[src] * L / 90.0
    # the calls' pointers are marshalled once, as a host model sets up its
    # radiation calls once and runs them every time step
    reqs = R.Requests(calls)
    res = {}
    for prec in ("fp64", "mixed"):
        b = R.Batch(prec)
        for _ in range(warmup):
            b.run(reqs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev_ms = 0.0
        for _ in range(steps):
            b.run(reqs)
            dev_ms += b.stats()["last_ms"]
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        st = b.stats()
        b.close()
        res[prec] = dict(ms_per_call=dt * 1e3, device_ms=dev_ms / steps,
                         solves_per_s_l90=solves / dt, captures=st["captures"], replays=st["replays"])
    line = {"benchmark": "ragged batch, one graph-captured launch per radiation call",
            "columns": n_columns, "calls_per_batch": len(calls), "layers": "U{30..137}",
            "split": "50% LW (9 bands, 50 g) / 50% SW (6 bands, 40 g), all-sky McICA",
            "fp64": res["fp64"], "mixed": res["mixed"],
            "mixed_over_fp64": res["mixed"]["solves_per_s_l90"] / res["fp64"]["solves_per_s_l90"]}
    if not quiet:
        print(json.dumps(line), flush=True)
    return line


def serve(rates=(50, 200, 1000, 2000), seconds=10.0, pool=48, seed=20261207, precisions=("fp64", "mixed"),
          max_requests=4000, inputs="host"):
    """inputs="host": each call hands over host arrays, which cross PCIe in
    and out (a host model whose state lives in host memory). inputs="device":
    the profiles already sit in HBM and each call writes fresh device output
    arrays (a host model whose state is device-resident)."""
    on_dev = inputs == "device"
    if on_dev:   # torch takes the device before the library does
        import torch
        dev = torch.device("cuda:0")
        torch.zeros(1, device=dev)
    from socrates_amd.serve import Server
    R.init(0)
    R.register_default_spectra()
    rng = np.random.default_rng(seed)
    templates = []
    for k in range(pool):
        src = "lw" if k % 2 == 0 else "sw"
        P = int(np.exp(rng.uniform(0.0, np.log(4096.0))))
        L = int(rng.integers(40, 138))
        case = synth.make_case(max(P, 1), L, src, NBAND[src], seed=seed + k, cloudy=True,
                               first_column_id=100000 * k)
        templates.append((src, case, _device_case(torch, case, dev)[0] if on_dev else None))

    def call(src, case, arrs):
        if arrs is None:
            return R.case_call(case, src)
        L, P = case.n_layer, case.n_profile
        out = {k: torch.empty(((L + 1) if k != "heating_rate" else L, P), dtype=torch.float64, device=dev)
               for k in ("flux_up", "flux_down", "flux_direct", "heating_rate")}
        return R.case_call(case, src, device_arrays=arrs, out=out)
    lines = []
    for prec in precisions:
        for rate in rates:
            srv = Server(precision=prec, hbm_budget_gb=96.0, n_slots=2, max_batch_requests=512)
            # warm the slots' buffers
            srv.wait(srv.submit(*call(*templates[0])))
            srv.reset_stats()
            n = int(min(max_requests, rate * seconds))
            gaps = rng.exponential(1.0 / rate, n)
            picks = rng.integers(0, pool, n)
            solves = 0.0
            # a consumer waits on the tickets as they come, releasing the
            # outputs of finished calls
            import queue
            q = queue.Queue()

            def consume():
                while True:
                    t = q.get()
                    if t is None:
                        return
                    srv.wait(t, 600.0)
            th = threading.Thread(target=consume, daemon=True)
            th.start()
            t_start = time.perf_counter()
            t_next = t_start
            for i in range(n):
                t_next += gaps[i]
                now = time.perf_counter()
                if t_next > now:
                    time.sleep(t_next - now)
                src, case, arrs = templates[picks[i]]
                q.put(srv.submit(*call(src, case, arrs)))
                solves += case.n_profile * NG[src] * case.n_layer / 90.0
            q.put(None)
            th.join()
            wall = time.perf_counter() - t_start
            st = srv.stats()
            srv.close()
            line = {"benchmark": "serving loop, Poisson arrivals", "inputs": inputs, "precision": prec,
                    "rate_req_s": rate,
                    "requests": n, "offered_s": float(gaps.sum()), "wall_s": wall,
                    "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "max_ms": st["max_ms"],
                    "mean_batch": st["mean_batch"], "batches": st["batches"],
                    "solves_per_s_l90": solves / wall}
            print(json.dumps(line), flush=True)
            lines.append(line)
    return lines


def triple(steps=10, warmup=2, P=4096, L=90, seed=20261209):
    """Stratiform + convective cloud as triple-column regions
    (solver_triple_hogan), SW and LW on P x L columns, one batch per step,
    FP64 and mixed; (column, k-term) solves per second."""
    import torch
    dev = torch.device("cuda:0")
    R.init(0)
    R.register_default_spectra()
    calls, keep = [], []
    for src in ("sw", "lw"):
        case = synth.make_case(P, L, src, NBAND[src], seed=seed + (src == "sw"), cloudy=True, overlap="triple")
        arrs, out = _device_case(torch, case, dev)
        for k in ("conv_frac", "liq_conv_mmr", "liq_conv_dim"):
            arrs[k] = torch.from_numpy(np.ascontiguousarray(getattr(case, k))).to(dev)
        keep.append((arrs, out))
        calls.append(R.case_call(case, src, device_arrays=arrs, out=out))
    reqs = R.Requests(calls)
    solves = P * (NG["sw"] + NG["lw"])
    res = {}
    for prec in ("fp64", "mixed"):
        b = R.Batch(prec)
        for _ in range(warmup):
            b.run(reqs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run(reqs)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        b.close()
        res[prec] = dict(ms_per_step=dt * 1e3, solves_per_s=solves / dt)
    line = {"benchmark": "triple-column cloud, one graph-captured batch of the SW and LW calls per step",
            "columns": P, "layers": L, "fp64": res["fp64"], "mixed": res["mixed"],
            "mixed_over_fp64": res["mixed"]["solves_per_s"] / res["fp64"]["solves_per_s"]}
    print(json.dumps(line), flush=True)
    return line


def ga9(steps=10, warmup=2, P=4096, L=90, seed=20261211):
    """The production GA9 spectral files (41 SW and 81 LW k-terms, P/T
    look-up tables, water-vapour continuum, ten further gases) with their
    McICA data on P x L all-sky columns, SW and LW in one graph-captured
    batch per step, FP64 and mixed; (column, k-term) solves per second."""
    import torch
    dev = torch.device("cuda:0")
    R.init(0)
    ga9_dir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "ga9")
    R.set_spectrum("sw_ga9", os.path.join(ga9_dir, "sp_sw_ga9"))
    R.set_spectrum("lw_ga9", os.path.join(ga9_dir, "sp_lw_ga9"))
    R.set_mcica(os.path.join(ga9_dir, "mcica_data"), sw_spectrum_name="sw_ga9", lw_spectrum_name="lw_ga9")
    wm = np.array([4.7255e-7, 9.637e-7, 0.23139, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    calls, keep = [], []
    n_g = {"sw": 41, "lw": 81}
    for src in ("sw", "lw"):
        case = synth.make_case(P, L, src, 6 if src == "sw" else 9, seed=seed + (src == "sw"), cloudy=True)
        case.wm_mix_ratio = wm.copy()
        arrs, out = _device_case(torch, case, dev)
        keep.append((arrs, out))
        calls.append(R.case_call(case, src + "_ga9", device_arrays=arrs, out=out))
    reqs = R.Requests(calls)
    solves = P * (n_g["sw"] + n_g["lw"])
    res = {}
    for prec in ("fp64", "mixed"):
        b = R.Batch(prec)
        for _ in range(warmup):
            b.run(reqs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run(reqs)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        b.close()
        res[prec] = dict(ms_per_step=dt * 1e3, solves_per_s=solves / dt)
    line = {"benchmark": "GA9 spectral files, all-sky McICA, one graph-captured batch of the SW and LW calls per step",
            "columns": P, "layers": L, "k_terms": n_g, "fp64": res["fp64"], "mixed": res["mixed"],
            "mixed_over_fp64": res["mixed"]["solves_per_s"] / res["fp64"]["solves_per_s"]}
    print(json.dumps(line), flush=True)
    return line


def aerosol(steps=10, warmup=2, P=4096, L=90, n_mode=3, seed=20261212):
    """The C2 workload (synthetic files, 4096 x 90 all-sky McICA, SW and LW
    in one graph-captured batch per step) with n_mode modal aerosols on
    every column, FP64 and mixed; and the same batch without them, so the
    cost of the aerosols (their sums in the band kernel, clear air with a
    phase function in the solve, the general build) is the difference."""
    import torch
    dev = torch.device("cuda:0")
    R.init(0)
    R.register_default_spectra()
    n_g = {"sw": 40, "lw": 50}
    res = {}
    for with_aer in (False, True):
        calls, keep = [], []
        for src in ("sw", "lw"):
            case = synth.make_case(P, L, src, 6 if src == "sw" else 9, seed=seed + (src == "sw"), cloudy=True)
            if with_aer:
                synth.add_aerosols(case, n_mode, seed + 10 + (src == "sw"))
            arrs, out = _device_case(torch, case, dev)
            if with_aer:
                for k in ("aer_mix_ratio", "aer_absorption", "aer_scattering", "aer_asymmetry"):
                    arrs[k] = torch.from_numpy(np.ascontiguousarray(getattr(case, k))).to(dev)
            keep.append((arrs, out))
            calls.append(R.case_call(case, src, device_arrays=arrs, out=out))
        reqs = R.Requests(calls)
        solves = P * (n_g["sw"] + n_g["lw"])
        for prec in ("fp64", "mixed"):
            b = R.Batch(prec)
            for _ in range(warmup):
                b.run(reqs)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                b.run(reqs)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / steps
            b.close()
            res[("aerosols" if with_aer else "none") + "_" + prec] = dict(ms_per_step=dt * 1e3,
                                                                         solves_per_s=solves / dt)
    line = {"benchmark": "C2 workload with %d modal aerosol modes vs none, one graph-captured batch of the "
                         "SW and LW calls per step" % n_mode, "columns": P, "layers": L, **res}
    print(json.dumps(line), flush=True)
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["ragged", "serve", "triple", "ga9", "aerosol"])
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--rates", default="50,200,1000,2000")
    ap.add_argument("--seconds", type=float, default=10.0)
    ap.add_argument("--inputs", choices=["host", "device"], default="host")
    ap.add_argument("--max-requests", type=int, default=4000)
    a = ap.parse_args()
    if a.what == "ragged":
        ragged(steps=a.steps)
    elif a.what == "triple":
        triple(steps=a.steps)
    elif a.what == "ga9":
        ga9(steps=a.steps)
    elif a.what == "aerosol":
        aerosol(steps=a.steps)
    else:
        serve(rates=[int(x) for x in a.rates.split(",")], seconds=a.seconds, max_requests=a.max_requests,
              inputs=a.inputs)


if __name__ == "__main__":
    main()
