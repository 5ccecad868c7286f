"""The GA9 spectral files on the GPU, through the C ABI: look-up-table,
power-quadratic and Wenyi k-terms, the water-vapour continuum and the
further well-mixed gases, clear sky, triple-column cloud and McICA.

FP64 form: the golden vectors of the reference compiled from its sources
and the oracle (pinned to them to rounding, tests/test_ga9.py), with the
tolerances of the McICA path (test_gpu_parity.check_fp64); the runes
example's known-good output to its 8 printed decimals. Mixed form: 0.05
W m-2 and 0.01 K/day of the FP64 oracle.
"""
import glob
import os

import numpy as np
import pytest

from ga9_cases import MCICA_GA9, SPEC_GA9, example_table, kgo, runes_example_case
from oracle import oracle
from socrates_amd import runes as R
from socrates_amd import synth
from socrates_amd.case_io import case_from_arrays
from test_gpu_parity import check_fp64, check_mixed

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAME = {"sw": "sw_ga9", "lw": "lw_ga9"}
WM = np.array([4.7255e-7, 9.637e-7, 0.23139, 2.0e-10, 8.0e-10, 1.1e-9, 2.4e-9, 3.3e-10, 9.0e-10,
               4.5e-10])


@pytest.fixture(scope="module", autouse=True)
def spectra():
    R.init(0)
    for src in ("sw", "lw"):
        R.set_spectrum(NAME[src], SPEC_GA9[src])
    R.set_mcica(MCICA_GA9, sw_spectrum_name=NAME["sw"], lw_spectrum_name=NAME["lw"])


def ga9_case(src, kind, P, L, seed):
    c = synth.make_case(P, L, src, 6 if src == "sw" else 9, seed=seed, cloudy=kind != "clear",
                        overlap="triple" if kind == "triple" else "mcica")
    c.wm_mix_ratio = WM.copy()
    return c


@pytest.mark.parametrize("src", ["sw", "lw", "lw2"])
def test_runes_example_on_gpu(src):
    """The example's three sections, the third with the surface emission
    per band (flux_ground)."""
    out = R.runes_case(runes_example_case(src), NAME["sw" if src == "sw" else "lw"], precision="fp64")
    err = np.abs(example_table(out) - kgo(src))
    assert err.max() <= 6e-9, err.max(axis=0)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ga9_runes_*.npz"))))
def test_ga9_golden_fp64(path):
    d = np.load(path)
    case = case_from_arrays(d)
    src = "sw" if case.i_source == 1 else "lw"
    ref = {k: d["ref_" + k] for k in ("flux_up", "flux_down", "flux_direct", "heating_rate")}
    gpu = R.runes_case(case, NAME[src], precision="fp64")
    check_fp64(gpu, ref, os.path.basename(path), src, case.layer_heat_capacity)


@pytest.mark.parametrize("src", ["sw", "lw"])
@pytest.mark.parametrize("kind", ["clear", "triple", "mcica"])
def test_ga9_fp64_matches_oracle(src, kind):
    case = ga9_case(src, kind, 40, 90, 1200 + len(kind) + (src == "sw"))
    ref = oracle.runes(case, SPEC_GA9[src], MCICA_GA9 if kind == "mcica" else None)
    gpu = R.runes_case(case, NAME[src], precision="fp64")
    check_fp64(gpu, ref, "ga9 %s %s" % (src, kind), src, case.layer_heat_capacity)


@pytest.mark.parametrize("src", ["sw", "lw"])
@pytest.mark.parametrize("kind", ["clear", "mcica"])
def test_ga9_mixed_within_bounds(src, kind):
    case = ga9_case(src, kind, 128, 90, 1300 + len(kind) + (src == "sw"))
    ref = oracle.runes(case, SPEC_GA9[src], MCICA_GA9 if kind == "mcica" else None)
    gpu = R.runes_case(case, NAME[src], precision="mixed")
    check_mixed(gpu, ref, "ga9 mixed %s %s" % (src, kind))


def test_ga9_continuum_needs_density():
    case = ga9_case("lw", "clear", 4, 30, 7)
    case.density = None
    with pytest.raises((R.SocgpuError, TypeError, AttributeError)):
        R.runes_case(case, NAME["lw"], precision="fp64")


def test_ga9_and_synthetic_calls_in_one_batch():
    """GA9 and synthetic-file calls in one batch each give the results they
    give alone, bit for bit."""
    R.register_default_spectra()
    calls = [(NAME["lw"], ga9_case("lw", "clear", 11, 60, 21)),
             ("sw", synth.make_case(5, 70, "sw", 6, seed=22, cloudy=True)),
             (NAME["sw"], ga9_case("sw", "triple", 9, 40, 23)),
             (NAME["sw"], ga9_case("sw", "mcica", 13, 50, 24))]
    for prec in ("fp64", "mixed"):
        alone = [R.runes_case(c, n, precision=prec) for n, c in calls]
        outs = R.runes_batch([R.case_call(c, n) for n, c in calls], precision=prec)
        for a, b in zip(alone, outs):
            for k in ("flux_up", "flux_down", "flux_direct", "heating_rate"):
                assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (prec, k)
