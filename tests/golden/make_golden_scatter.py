#!/usr/bin/env python3
"""Golden vectors for runes' treatment of scattering (i_scatter_method,
socrates_runes.F90:589-591; socrates_set_control.F90:272-277, 312-318,
368-370): synthetic infra-red columns on the GA9 files through the
reference compiled from its own sources (oracle/_ref/soc_runes_case),
clear sky and McICA, with scattering ignored (ip_scatter_none, absorption
only), extinction taken as absorption, approximate scattering and the
per-k-term choice of the spectral file (hybrid, the GA9 LW file marks its
terms), plus approximate scattering in the solar region. Writes
tests/golden/scatter_<src>_<kind>_m<method>.npz.

    python3 tests/golden/make_golden_scatter.py     (needs oracle/_ref built)
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
WM = np.array([4.7255e-7, 9.637e-7, 0.23139, 2.0e-10, 8.0e-10, 1.1e-9, 2.4e-9, 3.3e-10,
               9.0e-10, 4.5e-10])
# (source, kind, method, profiles, layers, seed)
CASES = [("lw", "clear", 2, 7, 50, 1201), ("lw", "clear", 3, 7, 50, 1202),
         ("lw", "clear", 5, 7, 50, 1203), ("lw", "mcica", 2, 6, 45, 1204),
         ("lw", "mcica", 3, 6, 45, 1205), ("lw", "mcica", 5, 6, 45, 1206),
         ("lw", "mcica", 4, 6, 45, 1207), ("sw", "mcica", 4, 6, 45, 1208)]


def scatter_case(src, kind, method, P, L, seed):
    nb = 6 if src == "sw" else 9
    c = synth.make_case(P, L, src, nb, seed=seed, cloudy=kind != "clear", overlap="mcica")
    c.wm_mix_ratio = WM.copy()
    c.i_scatter_method = method
    return c


def main():
    for src, kind, m, P, L, seed in CASES:
        c = scatter_case(src, kind, m, P, L, seed)
        mcica = os.path.join(GA9, "mcica_data") if kind == "mcica" else "-"
        with tempfile.TemporaryDirectory() as t:
            c.write(os.path.join(t, "case"))
            subprocess.run([REF, os.path.join(GA9, "sp_sw_ga9"), os.path.join(GA9, "sp_lw_ga9"), mcica,
                            os.path.join(t, "case"), os.path.join(t, "out")], check=True,
                           capture_output=True)
            ref = read_output(os.path.join(t, "out"), c.n_profile, c.n_layer)
        np.savez_compressed(os.path.join(HERE, "scatter_%s_%s_m%d.npz" % (src, kind, m)),
                            **case_to_arrays(c), **{"ref_" + k: v for k, v in ref.items()})


if __name__ == "__main__":
    main()
