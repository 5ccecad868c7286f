"""Triple-column cloud (stratiform and convective regions, maximum-random
overlap, solver_triple_hogan) on the GPU against the oracle, through the C
ABI.

FP64 form: the tolerances of the McICA path (test_gpu_parity.check_fp64).
Mixed form:
0.05 W m-2 and 0.01 K/day of the FP64 oracle. The triple-column oracle is
pinned bit-for-bit to the reference Fortran (tests/test_oracle.py).
"""
import os

import numpy as np
import pytest

from oracle import oracle
from socrates_amd import runes as R
from socrates_amd import synth
from test_gpu_parity import NBAND, SPEC, check_fp64, check_mixed

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def spectra():
    R.register_default_spectra()


def triple_case(P, L, src, seed):
    return synth.make_case(P, L, src, NBAND[src], seed=seed, cloudy=True, overlap="triple")


@pytest.mark.parametrize("src", ["lw", "sw"])
@pytest.mark.parametrize("shape", [(16, 60), (5, 137), (70, 30)])
def test_triple_fp64_matches_oracle(src, shape):
    P, L = shape
    case = triple_case(P, L, src, 300 + P + L + (src == "sw"))
    ref = oracle.runes(case, SPEC[src], None)
    gpu = R.runes_case(case, src, precision="fp64")
    check_fp64(gpu, ref, "triple %s %s" % (src, shape), src, case.layer_heat_capacity)


@pytest.mark.parametrize("src", ["lw", "sw"])
@pytest.mark.parametrize("where", ["top", "bottom", "none"])
def test_triple_cloud_top_cases(src, where):
    """Cloud in the top layer (cloud top 1), only in the bottom layer (cloud
    top L, where the reference couples the clear layer above without the
    overlap coefficients) and nowhere (cloud top L, all regions but the
    clear one empty)."""
    P, L = 9, 40
    case = triple_case(P, L, src, 77 + (src == "sw"))
    keep = np.zeros(L, bool)
    if where == "top":
        case.cloud_frac[0] = 0.4
        case.conv_frac[0] = 0.2
        case.liq_mmr[0] = 1e-4
        case.liq_conv_mmr[0] = 2e-4
    elif where == "bottom":
        keep[-1] = True
        case.cloud_frac[-1] = 0.3
        case.conv_frac[-1] = 0.1
        case.liq_mmr[-1] = 1e-4
        case.liq_conv_mmr[-1] = 1e-4
        case.cloud_frac[~keep] = 0.0
        case.conv_frac[~keep] = 0.0
    else:
        case.cloud_frac[:] = 0.0
        case.conv_frac[:] = 0.0
    ref = oracle.runes(case, SPEC[src], None)
    gpu = R.runes_case(case, src, precision="fp64")
    check_fp64(gpu, ref, "triple %s cloud %s" % (src, where), src, case.layer_heat_capacity)


@pytest.mark.parametrize("src", ["lw", "sw"])
def test_triple_mixed_within_bounds(src):
    P, L = 128, 90
    case = triple_case(P, L, src, 500 + (src == "sw"))
    ref = oracle.runes(case, SPEC[src], None)
    gpu = R.runes_case(case, src, precision="mixed")
    check_mixed(gpu, ref, "triple mixed %s" % src)


def test_triple_in_a_batch_with_mcica_calls():
    """A batch mixing triple-column, McICA and clear calls gives each call
    the results it gets alone, bit for bit."""
    cases = [("lw", triple_case(12, 50, "lw", 1)),
             ("sw", synth.make_case(7, 70, "sw", 6, seed=2, cloudy=True)),
             ("sw", triple_case(20, 33, "sw", 3)),
             ("lw", synth.make_case(5, 45, "lw", 9, seed=4, cloudy=False))]
    for prec in ("fp64", "mixed"):
        alone = [R.runes_case(c, s, precision=prec) for s, c in cases]
        outs = R.runes_batch([R.case_call(c, s) for s, c in cases], precision=prec)
        for a, b in zip(alone, outs):
            for k in ("flux_up", "flux_down", "flux_direct", "heating_rate"):
                assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (prec, k)


def test_triple_rejects_unsupported_overlap():
    """Random overlap with convective and stratiform cloud is
    ip_cloud_mix_random, which the reference rejects for the triple solver."""
    case = triple_case(4, 30, "lw", 5)
    case.i_overlap = 1
    with pytest.raises(R.SocgpuError):
        R.runes_case(case, "lw", precision="fp64")
