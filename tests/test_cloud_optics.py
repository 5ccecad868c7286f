"""Cloud condensate optics of the oracle pinned against the reference's
golden vectors (tests/golden/make_golden_ice.py, make_golden_slingo.py):

* ice of the GA9 files' types 8 (Fu's phase-function fit), 11 (Baran) and
  12 (Pade-2), McICA and triple-column cloud, SW and LW, bit for bit to
  1e-13; Baran ice with no condensate (its fit is not zero there);
* Slingo-Schrecker droplets (parametrisation 1) on the synthetic files
  with the GA3 Slingo fits.
Each golden differs from the case without the component, so it pins it.
"""
import dataclasses
import glob
import os

import numpy as np
import pytest

from oracle import oracle
from socrates_amd.case_io import case_from_arrays

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
DATA = os.path.join(ROOT, "socrates_amd", "data")
GA9 = os.path.join(DATA, "ga9")
OUTS = ("flux_up", "flux_down", "flux_direct", "heating_rate")


def spec_for(path, src):
    if "slingo" in os.path.basename(path):
        return os.path.join(DATA, "sp_%s_slingo" % src)
    return os.path.join(GA9, "sp_%s_ga9" % src)


def mcica_for(path):
    if "_mcica" not in os.path.basename(path):
        return None
    return os.path.join(DATA, "mcica_synth") if "slingo" in path else os.path.join(GA9, "mcica_data")


GOLDENS = sorted(glob.glob(os.path.join(GOLD, "ice_ga9_*.npz")) + glob.glob(os.path.join(GOLD, "slingo_*.npz")))


@pytest.mark.parametrize("path", GOLDENS)
def test_oracle_cloud_optics_match_reference(path):
    d = np.load(path)
    case = case_from_arrays(d)
    src = "sw" if case.i_source == 1 else "lw"
    out = oracle.runes(case, spec_for(path, src), mcica_for(path))
    for k in OUTS:
        r = d["ref_" + k]
        assert (np.abs(out[k] - r) <= 1e-13 * np.maximum(np.abs(r), 1.0)).all(), (k, np.abs(out[k] - r).max())


@pytest.mark.parametrize("path", GOLDENS)
def test_component_matters(path):
    """Leaving the ice out (no ice component) or swapping Slingo for the
    synthetic Pade-2 droplets changes the fluxes."""
    case = case_from_arrays(np.load(path))
    src = "sw" if case.i_source == 1 else "lw"
    base = oracle.runes(case, spec_for(path, src), mcica_for(path))
    if "slingo" in path:
        other = oracle.runes(case, os.path.join(DATA, "sp_%s_synth" % src), mcica_for(path))
    else:
        other = oracle.runes(dataclasses.replace(case, i_st_ice=0), spec_for(path, src), mcica_for(path))
    diff = np.abs(other["flux_up"] - base["flux_up"]).max()
    assert diff > (1e-12 if "_zero" in path else 1e-3), diff
