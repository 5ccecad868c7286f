"""The serving loop (socgpu_serve_*): calls submitted while others are on
the GPU are admitted as device memory frees up, packed into batches and run
alongside the batches in flight; every call's outputs are bit-identical to
the same call made alone, and identical whether it ran alone in the server
or in a mix."""
import numpy as np
import pytest

from socrates_amd import runes as R
from socrates_amd import synth
from socrates_amd.serve import Server

pytestmark = pytest.mark.gpu

NBAND = {"lw": 9, "sw": 6}


@pytest.fixture(scope="module", autouse=True)
def spectra():
    R.register_default_spectra()


def _cases(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        src = "sw" if k % 3 == 0 else "lw"
        P = int(np.exp(rng.uniform(0.0, np.log(600.0))))
        L = int(rng.integers(40, 138))
        out.append((src, synth.make_case(max(P, 1), L, src, NBAND[src], seed=seed + k, cloudy=True,
                                         first_column_id=10000 * k)))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_serve_alone_equals_mix(precision):
    cases = _cases(14, 31337)
    ref = [R.runes_case(c, src, precision=precision) for src, c in cases]
    srv = Server(precision=precision, hbm_budget_gb=16.0, n_slots=2)
    try:
        x_src, x_case = cases[5]
        alone = srv.wait(srv.submit(*R.case_call(x_case, x_src)))
        tickets = [srv.submit(*R.case_call(c, src)) for src, c in cases]
        outs = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    for k in ("flux_up", "flux_down", "heating_rate"):
        assert np.array_equal(alone[k], outs[5][k]), k
        assert np.array_equal(alone[k], ref[5][k]), k
    for (src, c), r, o in zip(cases, ref, outs):
        for k in ("flux_up", "flux_down", "heating_rate"):
            assert np.array_equal(r[k], o[k]), (src, c.n_profile, c.n_layer, k)
    assert st["completed"] == len(cases) + 1
    assert st["batches"] < len(cases) + 1, st   # calls were batched together


def test_serve_admission_by_footprint():
    """A budget that holds about one call at a time still completes every
    call, in order of admission, with the same results."""
    cases = _cases(6, 99)
    ref = [R.runes_case(c, src) for src, c in cases]
    srv = Server(precision="fp64", hbm_budget_gb=0.25, n_slots=2)
    try:
        tickets = [srv.submit(*R.case_call(c, src)) for src, c in cases]
        outs = [srv.wait(t) for t in tickets]
    finally:
        srv.close()
    for r, o in zip(ref, outs):
        assert np.array_equal(r["flux_down"], o["flux_down"])


def test_serve_rejects_call_above_budget():
    src, c = _cases(1, 7)[0]
    big = synth.make_case(2000, 137, "lw", 9, seed=3, cloudy=True)
    srv = Server(precision="fp64", hbm_budget_gb=0.01, n_slots=1)
    try:
        with pytest.raises(R.SocgpuError):
            srv.submit(*R.case_call(big, "lw"))
    finally:
        srv.close()
