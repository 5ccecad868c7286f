#!/usr/bin/env python3
"""Golden vectors of the precision-stress spectral files
(socrates_amd/data/sp_{sw,lw}_stress, written by make_spectra.py) from the
reference compiled from its own sources (oracle/build_ref.py):

  stress_runes_<src>_<clear|mcica>.npz  inputs of one runes() call on the
                                        C1 column recipe (8 columns x 60
                                        layers) + the reference's fluxes and
                                        heating rates (K/s)

Runs in the build container only. Usage: python3 tests/golden/make_golden_stress.py
"""
import os
import resource
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
from socrates_amd import synth  # noqa: E402
from socrates_amd.case_io import case_to_arrays, read_output  # noqa: E402

REF = os.path.join(REPO, "oracle", "_ref")
DATA = os.path.join(REPO, "socrates_amd", "data")

CASES = {
    # name: (n_profile, n_layer, source, cloudy, seed)
    "sw_clear": (8, 60, "sw", False, 20261221),
    "sw_mcica": (8, 60, "sw", True, 20261222),
    "lw_clear": (8, 60, "lw", False, 20261223),
    "lw_mcica": (8, 60, "lw", True, 20261224),
}


def _stack():
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def main():
    with tempfile.TemporaryDirectory() as tmp:
        for name, (P, L, src, cloudy, seed) in CASES.items():
            case = synth.make_case(P, L, src, 9 if src == "lw" else 6, seed, cloudy=cloudy)
            fin, fout = os.path.join(tmp, name + ".bin"), os.path.join(tmp, name + ".out")
            case.write(fin)
            subprocess.run([os.path.join(REF, "soc_runes_case"), os.path.join(DATA, "sp_sw_stress"),
                            os.path.join(DATA, "sp_lw_stress"),
                            os.path.join(DATA, "mcica_synth") if cloudy else "-", fin, fout],
                           check=True, capture_output=True, preexec_fn=_stack)
            out = read_output(fout, P, L)
            np.savez_compressed(os.path.join(HERE, "stress_runes_%s.npz" % name), **case_to_arrays(case),
                                **{"ref_" + k: v for k, v in out.items()})
    print("stress golden vectors written")


if __name__ == "__main__":
    main()
