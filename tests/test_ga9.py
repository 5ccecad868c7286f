"""The oracle on the GA9 spectral files: look-up-table, power-quadratic and
Wenyi k-terms, the water-vapour continuum and the further well-mixed gases.

Pins:
- the reference's own known-good output of its runes example
  (examples/runes/gfortran_8_1_0.txt, 8 decimals) for the SW call and the
  first LW call;
- golden vectors from the reference compiled from its sources
  (tests/golden/make_golden_ga9.py): clear, triple-column and McICA columns
  with every further gas, to rounding.
"""
import glob
import os

import numpy as np
import pytest

from ga9_cases import GA9, MCICA_GA9, SPEC_GA9, example_table, kgo, runes_example_case
from oracle import oracle
from socrates_amd.case_io import case_from_arrays

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("src", ["sw", "lw", "lw2"])
def test_oracle_reproduces_runes_example(src):
    """All three sections of the example: SW, LW with the surface
    temperature, LW with the surface emission per band (flux_ground)."""
    out = oracle.runes(runes_example_case(src), SPEC_GA9["sw" if src == "sw" else "lw"])
    err = np.abs(example_table(out) - kgo(src))
    # printed to 8 decimals: half a unit of the last place, plus rounding
    assert err.max() <= 6e-9, err.max(axis=0)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ga9_runes_*.npz"))))
def test_oracle_matches_reference_golden_ga9(path):
    d = np.load(path)
    case = case_from_arrays(d)
    src = "sw" if case.i_source == 1 else "lw"
    mcica = MCICA_GA9 if case.l_cloud and case.i_cloud_repr == 1 else None
    out = oracle.runes(case, SPEC_GA9[src], mcica)
    for k in ("flux_up", "flux_down", "flux_direct", "heating_rate"):
        r = d["ref_" + k]
        err = np.abs(out[k] - r)
        assert (err <= 1e-13 * np.maximum(np.abs(r), 1.0)).all(), (k, err.max())


def test_golden_cases_carry_every_further_gas():
    """Each golden GA9 case gives every well-mixed gas the files list a
    non-zero amount, so a dropped or misrouted gas changes the result."""
    for path in glob.glob(os.path.join(GOLD, "ga9_runes_*.npz")):
        assert (np.load(path)["case_wm_mix_ratio"] > 0).all()


def test_flux_ground_section_differs_from_skin_temperature():
    """The example's surface emission per band is close to the Planck flux
    of its surface temperature, but the two LW KGO sections still differ by
    some 400 units of their last printed place, so reproducing the second
    to 6e-9 pins the flux_ground path (its division by the diffuse
    co-albedo is pinned by the forms goldens, the example's albedo being 0)."""
    assert np.abs(kgo("lw2")[:, 1] - kgo("lw")[:, 1]).max() > 1e-6


def test_continuum_and_further_gases_matter():
    """Dropping the continuum's gas (water) or a minor gas changes the
    fluxes: the restatement is sensitive to both terms the KGO pins."""
    base = oracle.runes(runes_example_case("lw"), SPEC_GA9["lw"])
    c = runes_example_case("lw")
    c.wm_mix_ratio = np.zeros(10)
    no_minor = oracle.runes(c, SPEC_GA9["lw"])
    assert np.abs(no_minor["flux_up"] - base["flux_up"]).max() > 0.1
    assert os.path.exists(os.path.join(GA9, "sp_lw_ga9_k"))


@pytest.mark.skipif(not os.path.isdir("/root/reference/src/radiance_core"),
                    reason="the reference tree is only present where the tables are generated")
def test_wenyi_tables_match_reference_source():
    """The committed Wenyi tables are what gen_wenyi.py reads from the
    reference's scale_wenyi.F90 (data, regenerated rather than edited)."""
    import subprocess
    import sys
    import tempfile
    here = os.path.dirname(os.path.abspath(__file__))
    gen = os.path.join(os.path.dirname(here), "socrates_amd", "csrc", "gen_wenyi.py")
    committed = open(os.path.join(os.path.dirname(gen), "wenyi_tables.h")).read()
    with tempfile.TemporaryDirectory() as t:
        # run a copy of the generator so that it writes into the temporary directory
        src = open(gen).read()
        g2 = os.path.join(t, "gen_wenyi.py")
        open(g2, "w").write(src)
        subprocess.run([sys.executable, g2], check=True, capture_output=True)
        assert open(os.path.join(t, "wenyi_tables.h")).read() == committed
