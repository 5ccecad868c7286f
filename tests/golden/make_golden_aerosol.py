#!/usr/bin/env python3
"""Golden vectors for modal aerosols in runes (n_aer_mode, aer_mix_ratio,
aer_absorption, aer_scattering, aer_asymmetry; socrates_runes.F90:133-135,
609-621; socrates_set_aer.F90:299-309; grey_opt_prop.F90:716-743 with
opt_prop_ukca_aerosol.F90): synthetic columns with three modes through the
reference compiled from its own sources (oracle/_ref/soc_runes_case), clear
sky, McICA, the triple column (stratiform and convective cloud) and the
mixed column, on the synthetic spectral files and on the GA9 files (the
infra-red GA9 case with the per-k-term scattering of the file, as LFRic
configures it). Writes tests/golden/aerosol_<files>_<src>_<kind>.npz.

    python3 tests/golden/make_golden_aerosol.py     (needs oracle/_ref built)
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

DATA = os.path.join(ROOT, "socrates_amd", "data")
GA9 = os.path.join(DATA, "ga9")
REF = os.path.join(ROOT, "oracle", "_ref", "soc_runes_case")
WM = np.array([4.7255e-7, 9.637e-7, 0.23139, 2.0e-10, 8.0e-10, 1.1e-9, 2.4e-9, 3.3e-10,
               9.0e-10, 4.5e-10])
# (files, source, kind, profiles, layers, seed, scattering method)
CASES = [("synth", "sw", "clear", 9, 60, 1301, 0), ("synth", "lw", "clear", 9, 60, 1302, 0),
         ("synth", "sw", "mcica", 8, 60, 1303, 0), ("synth", "lw", "mcica", 8, 60, 1304, 0),
         ("ga9", "sw", "mcica", 6, 50, 1305, 0), ("ga9", "lw", "mcica", 6, 50, 1306, 5),
         ("ga9", "lw", "clear", 6, 50, 1307, 0),
         ("synth", "sw", "triple", 7, 60, 1308, 0), ("synth", "lw", "triple", 7, 60, 1309, 0),
         ("synth", "sw", "mixcol", 7, 60, 1310, 0), ("ga9", "lw", "triple", 6, 50, 1311, 0)]


def files(kind_files):
    if kind_files == "ga9":
        return (os.path.join(GA9, "sp_sw_ga9"), os.path.join(GA9, "sp_lw_ga9"),
                os.path.join(GA9, "mcica_data"))
    return (os.path.join(DATA, "sp_sw_synth"), os.path.join(DATA, "sp_lw_synth"),
            os.path.join(DATA, "mcica_synth"))


def aerosol_case(fs, src, kind, P, L, seed, method):
    nb = {"synth": {"sw": 6, "lw": 9}, "ga9": {"sw": 6, "lw": 9}}[fs][src]
    c = synth.make_case(P, L, src, nb, seed=seed, cloudy=kind != "clear",
                        overlap="mcica" if kind in ("clear", "mcica") else "triple")
    if kind == "mixcol":
        # the mixed column: homogeneous cloud without McICA (maximum-random)
        c.conv_frac[:] = 0.0
        c.liq_conv_mmr[:] = 0.0
        c.i_cloud_repr, c.i_inhom = 1, 0
    if fs == "ga9":
        c.wm_mix_ratio = WM.copy()
    c.i_scatter_method = method
    return synth.add_aerosols(c, 3, seed + 50)


def main():
    for fs, src, kind, P, L, seed, method in CASES:
        c = aerosol_case(fs, src, kind, P, L, seed, method)
        sw, lw, mc = files(fs)
        with tempfile.TemporaryDirectory() as t:
            c.write(os.path.join(t, "case"))
            subprocess.run([REF, sw, lw, mc if kind == "mcica" else "-", os.path.join(t, "case"),
                            os.path.join(t, "out")], check=True, capture_output=True)
            ref = read_output(os.path.join(t, "out"), c.n_profile, c.n_layer)
        np.savez_compressed(os.path.join(HERE, "aerosol_%s_%s_%s.npz" % (fs, src, kind)),
                            **case_to_arrays(c), **{"ref_" + k: v for k, v in ref.items()})
        print(fs, src, kind, "flux_up[0] range", ref["flux_up"][0].min(), ref["flux_up"][0].max())


if __name__ == "__main__":
    main()
