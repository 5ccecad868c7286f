#!/usr/bin/env python3
"""Copy the reference's file-driver examples into tests/golden/replay: the
input fields (netCDF classic files, as the examples hold them) and the
known-good outputs (CDL text) of the runs the replay reproduces.

  mcica/     examples/netcdf/mcica (one 52-layer CRM column, GA7 files,
             McICA, exponential overlap, ice type 8; run_me: lw_ga7, sw_ga7)
  7460_28/   examples/netcdf/7460_28 (one 167-layer CRM column, GA9 files,
             maximum-random mixed column, ice type 8; run_me: lw_ga9_avg,
             sw_ga9_avg net fluxes; the solar run reads .surfsw2)
  circ6/     examples/netcdf/CIRC_case6 (one 69-layer CIRC column, GA7
             files; run_me: case6_lw_ga7_clr clear sky, case6_lw_ga7 mixed
             column with scattering ignored, -z 2)

Data only: files are copied byte for byte.

    python3 tests/golden/make_replay_fixtures.py [--ref /root/reference]
"""
import argparse
import os
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("t", "tl", "q", "o3", "co2", "ch4", "n2o", "o2", "cfc11", "cfc12", "clfr", "lwm", "iwm",
          "re", "ire", "tstar", "szen", "stoa", "surflw", "surfsw", "surfsw2")
CASES = {
    "mcica": ("examples/netcdf/mcica", "5118_142", ("ref_lw_ga7.hrts", "ref_sw_ga7.hrts")),
    "7460_28": ("examples/netcdf/7460_28", "7460_28", ("ref_lw_ga9_avg.nflx", "ref_sw_ga9_avg.nflx")),
    "circ6": ("examples/netcdf/CIRC_case6", "case6", ("ref_case6_lw_ga7_clr.hrts", "ref_case6_lw_ga7.hrts")),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    for case, (sub, base, kgos) in CASES.items():
        src = os.path.join(a.ref, sub)
        dst = os.path.join(HERE, "replay", case)
        os.makedirs(dst, exist_ok=True)
        for sfx in FIELDS:
            p = os.path.join(src, "%s.%s" % (base, sfx))
            if os.path.exists(p):
                shutil.copyfile(p, os.path.join(dst, "%s.%s" % (base, sfx)))
        for k in kgos:
            shutil.copyfile(os.path.join(src, k), os.path.join(dst, k))
        print(case, sorted(os.listdir(dst)))


if __name__ == "__main__":
    main()
