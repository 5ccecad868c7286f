"""Ice and Slingo-Schrecker cloud optics on the device against the
reference's golden vectors (tests/golden/ice_ga9_*.npz, slingo_*.npz;
make_golden_ice.py, make_golden_slingo.py): the FP64 form to the tolerances
of tests/test_gpu_parity.py (Baran ice with no condensate, whose effect is
far below them, bit for bit with the oracle's case without it is not
claimed; its reference vectors are matched like the others), the mixed form
within its bounds."""
import glob
import os

import numpy as np
import pytest

from socrates_amd import runes as R
from socrates_amd.case_io import case_from_arrays
from test_gpu_parity import check_fp64, check_mixed

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
DATA = os.path.join(ROOT, "socrates_amd", "data")
GA9 = os.path.join(DATA, "ga9")
OUTS = ("flux_up", "flux_down", "flux_direct", "heating_rate")


@pytest.fixture(scope="module", autouse=True)
def spectra():
    R.init(0)
    for src in ("sw", "lw"):
        R.set_spectrum("ga9_" + src, os.path.join(GA9, "sp_%s_ga9" % src))
        R.set_spectrum("slingo_" + src, os.path.join(DATA, "sp_%s_slingo" % src))
    R.set_mcica(os.path.join(GA9, "mcica_data"), sw_spectrum_name="ga9_sw", lw_spectrum_name="ga9_lw")
    R.set_mcica(os.path.join(DATA, "mcica_synth"), sw_spectrum_name="slingo_sw", lw_spectrum_name="slingo_lw")


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ice_ga9_*.npz")) +
                                        glob.glob(os.path.join(GOLD, "slingo_*.npz"))))
def test_cloud_optics_match_reference(path):
    d = np.load(path)
    case = case_from_arrays(d)
    src = "sw" if case.i_source == 1 else "lw"
    name = ("slingo_" if "slingo" in os.path.basename(path) else "ga9_") + src
    ref = {k: d["ref_" + k] for k in OUTS}
    label = os.path.basename(path)
    print("case", label, flush=True)
    out = R.runes_case(case, name, precision="fp64")
    check_fp64(out, ref, label, src, case.layer_heat_capacity)
    mixed = R.runes_case(case, name, precision="mixed")
    check_mixed(mixed, ref, label + " mixed")


def test_ice_needs_its_type():
    d = np.load(os.path.join(GOLD, "ice_ga9_sw_mcica_t8.npz"))
    case = case_from_arrays(d)
    name, kw = R.case_call(case, "ga9_sw")
    P, L, s = kw.pop("n_profile"), kw.pop("n_layer"), kw.pop("i_source")
    with pytest.raises(R.SocgpuError):
        R.runes(P, L, name, s, **dict(kw, i_st_ice=0))
    with pytest.raises(R.SocgpuError):
        R.runes(P, L, name, s, **dict(kw, i_st_ice=9))   # a type the file does not hold
