#!/usr/bin/env python3
"""Write tests/golden/kgo_runes_ga9.npz: the inputs of the reference's runes
test driver (src/aux/runes_driver.f90:23-143: one mid-latitude-summer
column, 32 layers, GA9 spectral files) and its known-good outputs
(examples/runes/{gfortran_8_1_0,ifort_19_0_0,ifx_2025_0_4}.txt: the clear SW
call, the clear LW call with the surface temperature and the clear LW call
with the surface emission per band (flux_ground, runes_driver.f90:117-119,
232-252), flux down, flux up and heating rate at levels/layers 1..32, 8
decimals). Data only: the arrays are read from the
driver's source text and the KGO files.

    python3 tests/golden/make_ga9_kgo.py [--ref /root/reference]
"""
import argparse
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def f90_array(src, name):
    m = re.search(r"::\s*%s\([^)]*\)\s*=\s*reshape\(\s*\(/(.*?)/\)" % name, src, re.S)
    body = m.group(1).replace("&", " ")
    return np.array([float(x) for x in body.split(",")])


def f90_scalar(src, name):
    m = re.search(r"::\s*%s(?:\([^)]*\))?\s*=\s*([-0-9.eE+]+)" % name, src)
    return float(m.group(1))


def kgo_sections(path):
    out, cur = [], None
    for ln in open(path):
        if ln.startswith(("SW:", "LW:")):
            cur = []
            out.append(cur)
        elif cur is not None and ln.strip() and re.match(r"^\s+[-0-9.]+\s+[-0-9.]+\s+[-0-9.]+\s*$", ln):
            cur.append([float(x) for x in ln.split()])
    return [np.array(c) for c in out]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    src = open(os.path.join(a.ref, "src/aux/runes_driver.f90")).read()
    d = {k: f90_array(src, k) for k in ("p_layer", "t_layer", "p_level", "t_level", "h2o", "o3",
                                         "flux_ground")}
    for k in ("t_ground", "co2_mix_ratio", "ch4_mix_ratio", "n2o_mix_ratio", "o2_mix_ratio",
              "solar_irrad", "grey_albedo_sw", "grey_albedo_lw"):
        d[k] = np.array(f90_scalar(src, k))
    for tag in ("gfortran_8_1_0", "ifort_19_0_0", "ifx_2025_0_4"):
        secs = kgo_sections(os.path.join(a.ref, "examples/runes", tag + ".txt"))
        d["kgo_sw_" + tag] = secs[0]
        d["kgo_lw_" + tag] = secs[1]
        d["kgo_lw2_" + tag] = secs[2]
    np.savez(os.path.join(HERE, "kgo_runes_ga9.npz"), **d)


if __name__ == "__main__":
    main()
