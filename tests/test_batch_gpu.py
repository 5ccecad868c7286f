"""Ragged batches (socgpu_runes_batch, socgpu_batch_*): several radiation
calls with different spectra, sources, column counts and layer counts in
one packed launch sequence, and the persistent graph-captured form.

Every call's outputs are bit-identical to the same call made alone: the
packing changes which waves do the work, not the arithmetic of a column.
"""
import numpy as np
import pytest

from socrates_amd import runes as R
from socrates_amd import synth

pytestmark = pytest.mark.gpu

NBAND = {"lw": 9, "sw": 6}


@pytest.fixture(scope="module", autouse=True)
def spectra():
    R.register_default_spectra()


def ragged_cases(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        src = "lw" if k % 2 == 0 else "sw"
        P = int(rng.integers(1, 40))
        L = int(rng.integers(30, 138))
        cloudy = bool(rng.integers(0, 4) > 0)
        out.append((src, synth.make_case(P, L, src, NBAND[src], seed=seed + 17 * k, cloudy=cloudy,
                                         first_column_id=1000 * k)))
    return out


def _alone(cases, precision):
    return [R.runes_case(c, src, precision=precision) for src, c in cases]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_calls_alone(precision):
    cases = ragged_cases(9, 4242)
    alone = _alone(cases, precision)
    mixed = R.runes_batch([R.case_call(c, src) for src, c in cases], precision=precision)
    for (src, c), a, b in zip(cases, alone, mixed):
        for k in ("flux_up", "flux_down", "heating_rate"):
            assert np.array_equal(a[k], b[k]), (src, c.n_profile, c.n_layer, k)
        if src == "sw":
            assert np.array_equal(a["flux_direct"], b["flux_direct"])


def test_graph_batch_replays_and_matches():
    cases = ragged_cases(7, 777)
    alone = _alone(cases, "fp64")
    b = R.Batch("fp64")
    try:
        for rep in range(3):
            outs = b.run([R.case_call(c, src) for src, c in cases])
            for (src, c), a, o in zip(cases, alone, outs):
                for k in ("flux_up", "flux_down", "heating_rate"):
                    assert np.array_equal(a[k], o[k]), (rep, src, k)
        # a smaller batch fits the captured graph: replay, no new capture
        small = cases[:3]
        outs = b.run([R.case_call(c, src) for src, c in small])
        for (src, c), a, o in zip(small, alone[:3], outs):
            assert np.array_equal(a["flux_down"], o["flux_down"])
        st = b.stats()
        assert st["captures"] == 1 and st["replays"] == 4, st
    finally:
        b.close()


def test_batch_rejects_bad_request():
    src, c = ragged_cases(1, 5)[0]
    name, kw = R.case_call(c, src)
    kw["n_layer"] = 500
    with pytest.raises(R.SocgpuError):
        R.runes_batch([(name, kw)])


def test_reregistering_spectra_keeps_results():
    """Registering the spectra and McICA data again (the condensate table is
    shared by the spectra it is associated with, and freed with its last
    user) leaves the library usable and the results unchanged."""
    c = synth.make_case(9, 44, "sw", NBAND["sw"], seed=12, cloudy=True)
    before = R.runes_case(c, "sw", precision="fp64")
    for _ in range(3):
        R.register_default_spectra()
    after = R.runes_case(c, "sw", precision="fp64")
    for k in ("flux_up", "flux_down", "flux_direct", "heating_rate"):
        assert np.array_equal(np.asarray(before[k]), np.asarray(after[k])), k
