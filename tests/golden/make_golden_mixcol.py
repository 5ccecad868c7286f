#!/usr/bin/env python3
"""Golden vectors of the mixed-column cloud representation (ip_cloud_homogen
without McICA: ip_cloud_mix_max, ip_cloud_mix_random, ip_cloud_part_corr with
solver_mix_direct_hogan) and of the triple-column representation with
exponential-random overlap (ip_cloud_part_corr_cnv), from the reference
compiled from its sources (oracle/_ref/soc_runes_case), on the synthetic
spectral files. Writes tests/golden/runes_<src>_mixcol_ov<n>.npz and
runes_<src>_triple_ov3.npz.

    python3 tests/golden/make_golden_mixcol.py     (needs oracle/_ref built)
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
REF = os.path.join(ROOT, "oracle", "_ref", "soc_runes_case")


def mixcol_case(src, ov, P, L, seed, conv=False):
    """Cloudy columns with the stratiform fields of the triple-column cases;
    conv=False: the homogeneous (mixed-column) representation."""
    c = synth.make_case(P, L, src, 9 if src == "lw" else 6, seed=seed, cloudy=True, overlap="triple")
    c.i_overlap = ov
    c.cloud_vertical_decorr = synth.DECORR_PA
    c.conv_vertical_decorr = 0.5 * synth.DECORR_PA
    if not conv:
        c.conv_frac[:] = 0.0
        c.liq_conv_mmr[:] = 0.0
        c.i_cloud_repr, c.i_inhom, c.i_solver_kind = 1, 0, 0
    return c


CASES = [("sw", 0, False), ("lw", 0, False), ("sw", 1, False), ("lw", 1, False),
         ("sw", 3, False), ("lw", 3, False), ("sw", 3, True), ("lw", 3, True)]


def main():
    for k, (src, ov, conv) in enumerate(CASES):
        c = mixcol_case(src, ov, 7, 55, 7100 + k, conv)
        with tempfile.TemporaryDirectory() as t:
            c.write(os.path.join(t, "case"))
            subprocess.run([REF, os.path.join(DATA, "sp_sw_synth"), os.path.join(DATA, "sp_lw_synth"), "-",
                            os.path.join(t, "case"), os.path.join(t, "out")], check=True, capture_output=True)
            ref = read_output(os.path.join(t, "out"), c.n_profile, c.n_layer)
        name = "runes_%s_%s_ov%d.npz" % (src, "triple" if conv else "mixcol", ov)
        np.savez_compressed(os.path.join(HERE, name), **case_to_arrays(c),
                            **{"ref_" + k2: v for k2, v in ref.items()})


if __name__ == "__main__":
    main()
