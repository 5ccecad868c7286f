#!/usr/bin/env python3
"""Slingo-Schrecker droplets (droplet parametrisation 1,
opt_prop_water_cloud.F90:231-247): spectral files socrates_amd/data/
sp_{sw,lw}_slingo, the synthetic files with their type-5 droplet block
replaced by a parametrisation-1 block whose fits are the reference's GA3
Slingo-Schrecker droplets (data/spectra/ga3_0/sp_sw_ga3_0 type 2,
sp_lw_ga3_0 type 1, data copied as data), and golden vectors of the
reference compiled from its own sources (oracle/_ref/soc_runes_case) on
them: tests/golden/slingo_<src>_<kind>.npz, McICA and triple-column cloud.

    python3 tests/golden/make_golden_slingo.py [--ref /root/reference]
"""
import argparse
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


def block(lines, start):
    """The lines of the block that starts at index start (to its *END)."""
    end = start
    while not lines[end].startswith("*END"):
        end += 1
    return end


def droplet_block(path, dtype):
    lines = open(path).read().split("\n")
    for i, ln in enumerate(lines):
        if ln.startswith("*BLOCK: TYPE =   10") and int(lines[i + 2].split("=")[1]) == dtype:
            return lines[i:block(lines, i) + 1]
    raise ValueError("no droplet type %d in %s" % (dtype, path))


def write_variant(src, dst, slingo):
    """src with every droplet block dropped and the Slingo block added as
    droplet type 5 (range widened to the synthetic radii)."""
    lines = open(src).read().split("\n")
    out, i = [], 0
    while i < len(lines):
        if lines[i].startswith("*BLOCK: TYPE =   10"):
            i = block(lines, i) + 1
            continue
        out.append(lines[i])
        i += 1
    s = list(slingo)
    s[2] = " Type number of droplets =     5"
    s[5] = "Range of validity of parametrization =  1.50000E-06 --  5.00000E-05"
    while out and out[-1] == "":
        out.pop()
    open(dst, "w").write("\n".join(out + s) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    ga3 = os.path.join(a.ref, "data", "spectra", "ga3_0")
    write_variant(os.path.join(DATA, "sp_sw_synth"), os.path.join(DATA, "sp_sw_slingo"),
                  droplet_block(os.path.join(ga3, "sp_sw_ga3_0"), 2))
    write_variant(os.path.join(DATA, "sp_lw_synth"), os.path.join(DATA, "sp_lw_slingo"),
                  droplet_block(os.path.join(ga3, "sp_lw_ga3_0"), 1))
    for src in ("sw", "lw"):
        for kind in ("mcica", "triple"):
            c = synth.make_case(7, 45, src, 6 if src == "sw" else 9, seed=1701 + (src == "sw"),
                                cloudy=True, overlap=kind)
            mc = os.path.join(DATA, "mcica_synth") if kind == "mcica" else "-"
            with tempfile.TemporaryDirectory() as t:
                c.write(os.path.join(t, "case"))
                subprocess.run([REF, os.path.join(DATA, "sp_sw_slingo"), os.path.join(DATA, "sp_lw_slingo"),
                                mc, os.path.join(t, "case"), os.path.join(t, "out")], check=True,
                               capture_output=True)
                ref = read_output(os.path.join(t, "out"), c.n_profile, c.n_layer)
            np.savez_compressed(os.path.join(HERE, "slingo_%s_%s.npz" % (src, kind)), **case_to_arrays(c),
                                **{"ref_" + k: v for k, v in ref.items()})


if __name__ == "__main__":
    main()
