#!/usr/bin/env python3
"""Golden vectors with ice condensate on the GA9 spectral files, written by
the reference compiled from its own sources (oracle/_ref/soc_runes_case):
synthetic cloudy columns with stratiform (and, for the triple column,
convective) ice of the files' ice types 8 (Fu's phase-function fit,
parametrisation 12), 11 (Baran, 19) and 12 (Pade-2, 20), McICA and
triple-column cloud, SW and LW; and Baran ice with no ice condensate (its
fit adds extinction at zero ice). Writes tests/golden/ice_ga9_<src>_<kind>_t<type>[_zero].npz.

    python3 tests/golden/make_golden_ice.py     (needs oracle/_ref built)
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


def ice_case(src, kind, ityp, seed, zero=False):
    rng = np.random.default_rng(seed)
    P, L = 6, 45
    c = synth.make_case(P, L, src, 6 if src == "sw" else 9, seed=seed, cloudy=True, overlap=kind)
    c.wm_mix_ratio = WM.copy()
    c.i_st_ice = ityp
    scale = 0.0 if zero else 1.0
    c.ice_mmr = scale * np.where(c.cloud_frac > 0, rng.uniform(1e-6, 2e-4, (L, P)), 0.0) * c.cloud_frac
    c.ice_dim = rng.uniform(1e-5, 1.2e-4, (L, P))
    if kind == "triple":
        c.ice_conv_mmr = scale * np.where(c.conv_frac > 0, rng.uniform(1e-6, 2e-4, (L, P)), 0.0) * c.conv_frac
        c.ice_conv_dim = rng.uniform(1e-5, 1.2e-4, (L, P))
    return c


def main():
    n = 0
    for src in ("sw", "lw"):
        for kind in ("mcica", "triple"):
            for ityp, zero in ((8, False), (11, False), (12, False), (11, True)):
                n += 1
                c = ice_case(src, kind, ityp, 1800 + n, zero)
                mc = os.path.join(GA9, "mcica_data") if kind == "mcica" else "-"
                with tempfile.TemporaryDirectory() as t:
                    c.write(os.path.join(t, "case"))
                    subprocess.run([REF, os.path.join(GA9, "sp_sw_ga9"), os.path.join(GA9, "sp_lw_ga9"), mc,
                                    os.path.join(t, "case"), os.path.join(t, "out")], check=True,
                                   capture_output=True)
                    ref = read_output(os.path.join(t, "out"), c.n_profile, c.n_layer)
                name = "ice_ga9_%s_%s_t%d%s.npz" % (src, kind, ityp, "_zero" if zero else "")
                np.savez_compressed(os.path.join(HERE, name), **case_to_arrays(c),
                                    **{"ref_" + k: v for k, v in ref.items()})


if __name__ == "__main__":
    main()
