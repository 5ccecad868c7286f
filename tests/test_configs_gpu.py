"""Every benchmark configuration at its stated workload, through the C ABI,
against the oracle (the C restatement, itself pinned bit-for-bit against
the reference compiled from its own sources).

  C1  FP64 clear-sky LW, 64 columns x 60 layers (the reference's own CPU
      case): FP64 form within the FP64 tolerances, mixed form within its
      bounds; measured errors printed.
  C3  precision stress on the C1 shapes through the hot path (the packed
      solve kernel): the stress spectral files give g-points with
      single-scattering co-albedo 10^U(-8,-1) and layer optical depths
      1e-6..1e3, cloudy McICA sub-columns with asymmetry 0..0.95; both forms
      against the FP64 oracle per output (up, down, net flux in W m-2,
      heating rate in K/day).
  C4  10 000 ragged columns (levels U{30..137}, LW/SW 50/50, all-sky McICA)
      in one graph-replayed batch with autotuned tiles; every call against
      the oracle, FP64 within the FP64 tolerances, mixed within its bounds.
  C5  serving: Poisson arrivals of calls with 1..4096 columns (log-uniform),
      40..137 levels, LW or SW, all-sky McICA; every result bit-identical to
      the same call made alone, and a sample against the oracle.

Tolerances (written here, as in test_gpu_parity): FP64 1e-10 relative,
floors 1e-9 W m-2 and 1e-11 of the column's largest flux, plus, for the
workloads of thousands of columns, the oracle's own movement under one-ulp
changes of exp (1x, measured element by element) and, for the
precision-stress files, 8x it (near-conservative scattering is
ill-conditioned in the reference's own formulas); mixed 0.05 W m-2 and
0.01 K/day.
"""
import os

import numpy as np
import pytest

from oracle import oracle
from oracle_pool import run_many
from socrates_amd import errstats
from socrates_amd import runes as R
from socrates_amd import synth
from test_gpu_parity import MCICA, NBAND, SPEC, check_fp64, check_mixed

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "socrates_amd", "data")
STRESS = {"sw": os.path.join(DATA, "sp_sw_stress"), "lw": os.path.join(DATA, "sp_lw_stress")}


@pytest.fixture(scope="module", autouse=True)
def spectra():
    R.register_default_spectra()
    R.set_spectrum("sw_stress", STRESS["sw"])
    R.set_spectrum("lw_stress", STRESS["lw"])
    R.set_mcica(MCICA, sw_spectrum_name="sw_stress", lw_spectrum_name="lw_stress")


def _spread(jobs_out, ref):
    sp = {k: np.zeros_like(ref[k]) for k in ("flux_up", "flux_down", "flux_direct")}
    for j in jobs_out:
        for k in sp:
            sp[k] = np.maximum(sp[k], np.abs(j[k] - ref[k]))
    return sp


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c1_lw_clear_64x60(precision):
    case = synth.make_case(64, 60, "lw", 9, seed=20261202 + 1001, cloudy=False)
    ref = oracle.runes(case, SPEC["lw"], None)
    gpu = R.runes_case(case, "lw", precision=precision)
    print(errstats.line("C1 64x60 LW clear %s" % precision, errstats.stats(gpu, ref)))
    if precision == "fp64":
        check_fp64(gpu, ref, "C1", "lw", case.layer_heat_capacity)
    else:
        check_mixed(gpu, ref, "C1 mixed")


@pytest.mark.parametrize("src", ["sw", "lw"])
@pytest.mark.parametrize("cloudy", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c3_precision_stress_64x60(src, cloudy, precision):
    case = synth.make_case(64, 60, src, NBAND[src], seed=20261202 + 3003 + cloudy * 7 + (src == "sw"),
                           cloudy=cloudy)
    ref = oracle.runes(case, STRESS[src], MCICA if cloudy else None)
    gpu = R.runes_case(case, src + "_stress", precision=precision)
    print(errstats.line("C3 stress %s cloudy=%s %s" % (src, cloudy, precision), errstats.stats(gpu, ref)))
    if precision == "fp64":
        # near-conservative scattering (omega -> 1) is ill-conditioned in the
        # reference's own formulas (1 - e^2 gamma^2 -> 0): the tolerance adds
        # the oracle's measured sensitivity to one-ulp changes of exp
        jit = [oracle.runes(case, STRESS[src], MCICA if cloudy else None, variant="jitter") for _ in range(2)]
        check_fp64(gpu, ref, "C3 %s cloudy=%s" % (src, cloudy), src, case.layer_heat_capacity,
                   spread=_spread(jit, ref), factor=8.0)
    else:
        check_mixed(gpu, ref, "C3 mixed %s cloudy=%s" % (src, cloudy))


def _c4_calls(n_columns=10000, seed=20261206):
    rng = np.random.default_rng(seed)
    levels = rng.integers(30, 138, n_columns)
    source = np.where(np.arange(n_columns) < n_columns // 2, "lw", "sw")
    out = []
    for src in ("lw", "sw"):
        for L in np.unique(levels[source == src]):
            P = int(((levels == L) & (source == src)).sum())
            case = synth.make_case(P, int(L), src, NBAND[src], seed=seed + 1000 * int(L) + (src == "sw"),
                                   cloudy=True, first_column_id=int(L) * 100000)
            out.append((src, case))
    return out


def test_c4_ragged_10000_columns():
    cases = _c4_calls()
    assert sum(c.n_profile for _, c in cases) == 10000
    jobs = []
    for src, c in cases:
        for variant in (None, "jitter", "jitter"):
            jobs.append((c, SPEC[src], MCICA, variant))
    res = run_many(jobs)
    calls = [R.case_call(c, src) for src, c in cases]
    for precision in ("fp64", "mixed"):
        tile = R.autotune(calls, precision)
        b = R.Batch(precision)
        try:
            b.run(calls)               # captures
            outs = b.run(calls)        # replays
        finally:
            b.close()
        worst = {}
        for k, ((src, c), o) in enumerate(zip(cases, outs)):
            ref = res[3 * k]
            st = errstats.stats(o, ref)
            for name, v in st.items():
                w = worst.setdefault(name, dict(max=0.0, max_rel=0.0))
                w["max"] = max(w["max"], v["max"])
                w["max_rel"] = max(w["max_rel"], v["max_rel"])
            label = "C4 %s %s P=%d L=%d" % (precision, src, c.n_profile, c.n_layer)
            if precision == "fp64":
                check_fp64(o, ref, label, src, c.layer_heat_capacity, spread=_spread(res[3 * k + 1:3 * k + 3], ref))
            else:
                check_mixed(o, ref, label)
        print("C4 10000 ragged columns %s, tile %s: %s" % (precision, tile, "; ".join(
            "%s max %.3e (rel %.2e)" % (n, w["max"], w["max_rel"]) for n, w in worst.items())))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c5_poisson_serving(precision):
    import time
    from socrates_amd.serve import Server
    rng = np.random.default_rng(20261208)
    pool = []
    for k in range(12):
        src = "lw" if k % 2 == 0 else "sw"
        n = max(1, int(np.exp(rng.uniform(0.0, np.log(4096.0)))))
        if k == 0:
            n = 4096     # the largest call of the range
        if k == 1:
            n = 1        # and the smallest
        L = int(rng.integers(40, 138))
        pool.append((src, synth.make_case(n, L, src, NBAND[src], seed=20261208 + k, cloudy=True,
                                          first_column_id=100000 * k)))
    alone = [R.runes_case(c, src, precision=precision) for src, c in pool]
    srv = Server(precision=precision, hbm_budget_gb=16.0, n_slots=2)
    rate = 1000.0
    n_req = 80
    picks = rng.integers(0, len(pool), n_req)
    gaps = rng.exponential(1.0 / rate, n_req)
    try:
        tickets = []
        t_next = time.perf_counter()
        for i in range(n_req):
            t_next += gaps[i]
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            src, c = pool[picks[i]]
            tickets.append(srv.submit(*R.case_call(c, src)))
        outs = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    for i, o in enumerate(outs):
        a = alone[picks[i]]
        for k in ("flux_up", "flux_down", "flux_direct", "heating_rate"):
            assert np.array_equal(np.asarray(o[k]), np.asarray(a[k])), (i, int(picks[i]), k)
    assert st["completed"] == n_req
    print("C5 %s: %d calls at %.0f req/s offered, %d batches (mean %.2f calls), p50 %.2f ms, p99 %.2f ms"
          % (precision, n_req, rate, st["batches"], st["mean_batch"], st["p50_ms"], st["p99_ms"]))
    # a sample against the oracle: the largest, the smallest and two others
    sample = [0, 1, 4, 7]
    refs = run_many([(pool[j][1], SPEC[pool[j][0]], MCICA, None) for j in sample])
    for j, ref in zip(sample, refs):
        src, c = pool[j]
        o = alone[j]
        print(errstats.line("C5 sample %s P=%d L=%d %s" % (src, c.n_profile, c.n_layer, precision),
                            errstats.stats(o, ref)))
        if precision == "fp64":
            jit = run_many([(c, SPEC[src], MCICA, "jitter")] * 2)
            check_fp64(o, ref, "C5 sample %d" % j, src, c.layer_heat_capacity, spread=_spread(jit, ref))
        else:
            check_mixed(o, ref, "C5 sample %d mixed" % j)
