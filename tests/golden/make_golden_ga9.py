#!/usr/bin/env python3
"""Golden vectors on the GA9 spectral files (socrates_amd/data/ga9, copied as
data from the reference's data/spectra/ga9): synthetic columns through the
reference compiled from its own sources (oracle/_ref/soc_runes_case), SW and
LW, clear sky, triple-column cloud and McICA, with every further well-mixed
gas the files list. Writes tests/golden/ga9_runes_<src>_<kind>.npz.

    python3 tests/golden/make_golden_ga9.py     (needs oracle/_ref built)
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from socrates_amd import synth  # noqa: E402
from socrates_amd.case_io import case_to_arrays, read_output  # noqa: E402

GA9 = os.path.join(ROOT, "socrates_amd", "data", "ga9")
REF = os.path.join(ROOT, "oracle", "_ref", "soc_runes_case")
# n2o, ch4, o2, so2, ocs, cfc11, cfc12, cfc113, hcfc22, hfc134a (kg/kg)
WM = np.array([4.7255e-7, 9.637e-7, 0.23139, 2.0e-10, 8.0e-10, 1.1e-9, 2.4e-9, 3.3e-10,
               9.0e-10, 4.5e-10])


def ga9_case(src, kind, P, L, seed):
    nb = 6 if src == "sw" else 9
    c = synth.make_case(P, L, src, nb, seed=seed, cloudy=kind != "clear",
                        overlap="triple" if kind == "triple" else "mcica")
    c.wm_mix_ratio = WM.copy()
    return c


CASES = [("sw", "clear", 9, 60, 901), ("lw", "clear", 9, 60, 902),
         ("sw", "triple", 7, 45, 903), ("lw", "triple", 7, 45, 904),
         ("sw", "mcica", 6, 50, 905), ("lw", "mcica", 6, 50, 906)]


def main():
    for src, kind, P, L, seed in CASES:
        c = ga9_case(src, kind, P, L, seed)
        mcica = os.path.join(GA9, "mcica_data") if kind == "mcica" else "-"
        with tempfile.TemporaryDirectory() as t:
            c.write(os.path.join(t, "case"))
            subprocess.run([REF, os.path.join(GA9, "sp_sw_ga9"), os.path.join(GA9, "sp_lw_ga9"), mcica,
                            os.path.join(t, "case"), os.path.join(t, "out")], check=True,
                           capture_output=True)
            ref = read_output(os.path.join(t, "out"), c.n_profile, c.n_layer)
        np.savez_compressed(os.path.join(HERE, "ga9_runes_%s_%s.npz" % (src, kind)), **case_to_arrays(c),
                            **{"ref_" + k: v for k, v in ref.items()})


if __name__ == "__main__":
    main()
