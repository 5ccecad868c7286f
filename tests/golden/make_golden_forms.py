#!/usr/bin/env python3
"""Golden vectors of runes' further input and output forms, written by the
reference compiled from its own sources (oracle/_ref):

* forms_ga9_<src>.npz: GA9 McICA columns with a CO2 field, per-layer N2O
  and CH4 (and O2 in the SW) fields, the LW's surface emission per band
  (l_flux_ground) and the clear-sky diagnostics flux_up_clear,
  flux_down_clear, flux_direct_clear (soc_runes_case ... clear);
* lfric_<src>_<kind>[_drop<ip>].npz: synthetic columns through the host
  program socrates_amd/fortran/runes_lfric_case.f90 built against the
  reference: a profile list over strided bottom-up fields with profiles
  last, gases as strided fields and well-mixed ratios with their switches,
  clear-sky, surface and top-of-atmosphere diagnostics; _drop<ip>: the
  spectra registered without absorber <ip> (set_spectrum's gas flags). The
  same case through the plain call is stored beside it (ref_*);
* lfric_<src>_<kind>_aer.npz: the same with three modal aerosol modes,
  stored as the host model stores them (aer_*_1d, profiles last,
  bottom-up).

    python3 tests/golden/make_golden_forms.py --aerosols   (only the _aer files)

    python3 tests/golden/make_golden_forms.py     (needs oracle/_ref built)
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
DATA = os.path.join(ROOT, "socrates_amd", "data")
REF = os.path.join(ROOT, "oracle", "_ref")
WM = np.array([4.7255e-7, 9.637e-7, 0.23139, 2.0e-10, 8.0e-10, 1.1e-9, 2.4e-9, 3.3e-10,
               9.0e-10, 4.5e-10])


def run(exe, sw, lw, mcica, case, extra=()):
    with tempfile.TemporaryDirectory() as t:
        case.write(os.path.join(t, "case"))
        subprocess.run([os.path.join(REF, exe), sw, lw, mcica or "-", os.path.join(t, "case"),
                        os.path.join(t, "out")] + list(extra), check=True, capture_output=True)
        return np.fromfile(os.path.join(t, "out"))


def forms_case(src, seed):
    rng = np.random.default_rng(seed)
    P, L = 6, 50
    nb = 6 if src == "sw" else 9
    c = synth.make_case(P, L, src, nb, seed=seed, cloudy=True, overlap="mcica")
    c.wm_mix_ratio = WM.copy()
    c.co2 = rng.uniform(3.5e-4, 9.0e-4, (L, P))
    c.co2[5, 1] = -1.0e-6       # clamped at zero
    c.wm_field = {0: WM[0] * rng.uniform(0.5, 2.0, (L, P)), 1: WM[1] * rng.uniform(0.5, 2.0, (L, P))}
    if src == "sw":
        c.wm_field[2] = WM[2] * rng.uniform(0.9, 1.1, (L, P))
    else:
        c.l_flux_ground = 1
        c.flux_ground = rng.uniform(2.0, 80.0, (nb, P))
    return c


def aerosol_fixtures():
    sw, lw = os.path.join(DATA, "sp_sw_synth"), os.path.join(DATA, "sp_lw_synth")
    for src, kind, seed in (("sw", "mcica", 1331), ("lw", "clear", 1332), ("lw", "mcica", 1333),
                            ("sw", "triple", 1334)):
        c = synth.make_case(5, 30, src, 6 if src == "sw" else 9, seed=seed, cloudy=kind != "clear",
                            overlap="triple" if kind == "triple" else "mcica")
        synth.add_aerosols(c, 3, seed + 5)
        mc = os.path.join(DATA, "mcica_synth") if kind == "mcica" else None
        raw = run("runes_lfric_case", sw, lw, mc, c)
        d = {"lfric_raw": raw, "drop": np.array(0)}
        a = run("soc_runes_case", sw, lw, mc, c, ["1", "clear"])
        with tempfile.TemporaryDirectory() as t:
            a.tofile(os.path.join(t, "o"))
            ref = read_output(os.path.join(t, "o"), c.n_profile, c.n_layer, clear=True)
        d.update({"ref_" + k: v for k, v in ref.items()})
        np.savez_compressed(os.path.join(HERE, "lfric_%s_%s_aer.npz" % (src, kind)), **case_to_arrays(c), **d)


def main():
    if "--aerosols" in sys.argv:
        aerosol_fixtures()
        return
    aerosol_fixtures()
    for src, seed in (("sw", 1301), ("lw", 1302)):
        c = forms_case(src, seed)
        a = run("soc_runes_case", os.path.join(GA9, "sp_sw_ga9"), os.path.join(GA9, "sp_lw_ga9"),
                os.path.join(GA9, "mcica_data"), c, ["1", "clear"])
        with tempfile.TemporaryDirectory() as t:
            a.tofile(os.path.join(t, "o"))
            ref = read_output(os.path.join(t, "o"), c.n_profile, c.n_layer, clear=True)
        np.savez_compressed(os.path.join(HERE, "forms_ga9_%s.npz" % src), **case_to_arrays(c),
                            **{"ref_" + k: v for k, v in ref.items()})
    rng = np.random.default_rng(1310)
    for src in ("sw", "lw"):
        for kind in ("clear", "mcica", "triple"):
            drops = [0, 3] if kind == "clear" else [0]
            for drop in drops:
                c = synth.make_case(5, 30, src, 6 if src == "sw" else 9, seed=1311 + len(kind),
                                    cloudy=kind != "clear", overlap=kind if kind != "clear" else "mcica")
                c.co2 = rng.uniform(3.5e-4, 9.0e-4, (30, 5))
                if src == "lw" and kind != "triple":
                    c.l_flux_ground = 1
                    c.flux_ground = rng.uniform(2.0, 80.0, (c.n_band, 5))
                mc = os.path.join(DATA, "mcica_synth") if kind == "mcica" else None
                sw, lw = os.path.join(DATA, "sp_sw_synth"), os.path.join(DATA, "sp_lw_synth")
                raw = run("runes_lfric_case", sw, lw, mc, c, [str(drop)] if drop else [])
                d = {"lfric_raw": raw, "drop": np.array(drop)}
                if not drop:
                    a = run("soc_runes_case", sw, lw, mc, c, ["1", "clear"])
                    with tempfile.TemporaryDirectory() as t:
                        a.tofile(os.path.join(t, "o"))
                        ref = read_output(os.path.join(t, "o"), c.n_profile, c.n_layer, clear=True)
                    d.update({"ref_" + k: v for k, v in ref.items()})
                name = "lfric_%s_%s%s.npz" % (src, kind, "_drop%d" % drop if drop else "")
                np.savez_compressed(os.path.join(HERE, name), **case_to_arrays(c), **d)


if __name__ == "__main__":
    main()
