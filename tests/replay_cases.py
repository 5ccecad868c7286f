"""The reference's file-driver runs the replay reproduces: the examples'
input fields and known-good CDL outputs in tests/golden/replay (written by
tests/golden/make_replay_fixtures.py), with the Cl_run_cdf options of each
run as the examples' run_me scripts give them
(examples/netcdf/mcica/run_me:27-42, examples/netcdf/7460_28/run_me:33-47,
examples/netcdf/CIRC_case6/run_me:17-20)."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "replay")
DATA = os.path.join(os.path.dirname(HERE), "socrates_amd", "data")

# name: (folder, base name, surface file, known-good output, options)
RUNS = {
    "mcica_lw_ga7": ("mcica", "5118_142", "surflw", "ref_lw_ga7.hrts",
                     "-s {D}/ga7/sp_lw_ga7 -R 1 9 -I +R -g 4 -c -C 10 -K 1 -d 5 -i 8 -m 0.75 +m 2 -dp 10000"),
    "mcica_sw_ga7": ("mcica", "5118_142", "surfsw", "ref_sw_ga7.hrts",
                     "-s {D}/ga7/sp_sw_ga7 -R 1 6 -S +R -g 4 -r -C 10 -K 1 -d 5 -i 8 -m 0.75 +m 2 -dp 10000"),
    "crm_lw_ga9": ("7460_28", "7460_28", "surflw", "ref_lw_ga9_avg.nflx",
                   "-s {D}/ga9/sp_lw_ga9 -R 1 9 -I +R -g 4 -c -C 2 -K 1 -d 5 -i 8"),
    "crm_sw_ga9": ("7460_28", "7460_28", "surfsw2", "ref_sw_ga9_avg.nflx",
                   "-s {D}/ga9/sp_sw_ga9 -R 1 6 -S +R -g 4 -c -r -C 2 -K 1 -d 5 -i 8"),
    "circ6_lw_ga7_clr": ("circ6", "case6", "surflw", "ref_case6_lw_ga7_clr.hrts",
                         "-s {D}/ga7/sp_lw_ga7 -R 1 9 -I -g 4 -c -C 5"),
}


def run_args(name):
    folder, base, surf, kgo, opts = RUNS[name]
    return (os.path.join(GOLD, folder, base), opts.format(D=DATA).split() + ["+R"], surf,
            os.path.join(GOLD, folder, kgo))


def kgo_tolerance(vals, ulp):
    """Half a unit of the last printed digit (6 significant digits), plus
    the rounding of the value to single precision before it was printed
    (the output variables are REAL*4 in the file)."""
    return 0.5 * ulp + 2.0 ** -23 * np.abs(vals)
