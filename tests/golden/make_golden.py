#!/usr/bin/env python3
"""Write the golden vectors of the parity tests into tests/golden/*.npz.

Runs in the build container only: each case is computed by the reference
SOCRATES radiation core compiled from its own sources (oracle/build_ref.py,
harnesses in oracle/harness) and stored with its inputs:
  runes_<name>.npz      inputs of one runes() call + reference fluxes and
                        heating rates (K/s)
  cloud_gen_ov<k>.npz   inputs of cloud_gen + reference c_sub / n_subcol_cld
  two_stream_<v>.npz    prescribed optical properties of a kernel-level
                        two_stream call + reference direct and total fluxes
  meta.json             compiler, flags and seeds
Usage: python3 tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
from socrates_amd import synth  # noqa: E402
from socrates_amd.case_io import case_to_arrays, read_output  # noqa: E402

REF = os.path.join(REPO, "oracle", "_ref")
DATA = os.path.join(REPO, "socrates_amd", "data")

CASES = {
    # name: (n_profile, n_layer, source, cloudy, seed)
    "lw_clear": (8, 60, "lw", False, 20261203),
    "sw_clear": (8, 60, "sw", False, 20261204),
    "lw_mcica": (6, 60, "lw", True, 20261205),
    "sw_mcica": (6, 60, "sw", True, 20261206),
    "sw_mcica_l90": (4, 90, "sw", True, 20261207),
    "lw_triple": (6, 60, "lw", "triple", 20261212),
    "sw_triple": (6, 60, "sw", "triple", 20261213),
}


def run_case(name, P, L, src, cloudy, seed, tmp):
    overlap = "triple" if cloudy == "triple" else "mcica"
    case = synth.make_case(P, L, src, 9 if src == "lw" else 6, seed, cloudy=bool(cloudy), overlap=overlap)
    fin, fout = os.path.join(tmp, name + ".bin"), os.path.join(tmp, name + ".out")
    case.write(fin)
    subprocess.run([os.path.join(REF, "soc_runes_case"), os.path.join(DATA, "sp_sw_synth"),
                    os.path.join(DATA, "sp_lw_synth"),
                    os.path.join(DATA, "mcica_synth") if cloudy is True else "-", fin, fout],
                   check=True, capture_output=True)
    out = read_output(fout, P, L)
    np.savez_compressed(os.path.join(HERE, "runes_%s.npz" % name), **case_to_arrays(case),
                        **{"ref_" + k: v for k, v in out.items()})


def run_cloud_gen(ioverlap, tmp):
    rng = np.random.default_rng(4242 + ioverlap)
    P, L, S = 12, 40, 64
    st = synth.profiles(P, L, 99 + ioverlap)
    frac, _, _ = synth.clouds(st, 100 + ioverlap)
    frac = np.where(frac > 0.001, frac, 0.0)
    rsd = 0.75 + rng.uniform(-0.2, 0.2, (L, P))
    dp = np.full((L, P), synth.DECORR_PA)
    seeds = rng.integers(0, 200000, P).astype(np.int32)
    xcw = np.sort(np.abs(rng.normal(1.0, 0.6, (100, 40))), axis=0)
    fin, fout = os.path.join(tmp, "cg.bin"), os.path.join(tmp, "cg.out")
    with open(fin, "wb") as f:
        np.array([P, L, S, 1, ioverlap, 100, 40], dtype=np.int32).tofile(f)
        seeds.tofile(f)
        for a in (dp, dp * 0.5, rsd, frac, st["p_layer"]):
            np.ascontiguousarray(a).tofile(f)
        np.asfortranarray(xcw).ravel(order="F").tofile(f)
    subprocess.run([os.path.join(REF, "soc_cloud_gen_case"), fin, fout], check=True)
    raw = open(fout, "rb").read()
    ncld = np.frombuffer(raw, dtype=np.int32, count=P)
    c_sub = np.frombuffer(raw, dtype=np.float64, offset=4 * P).reshape(S, L, P)
    np.savez_compressed(os.path.join(HERE, "cloud_gen_ov%d.npz" % ioverlap), rand_seed=seeds,
                        rlc_cf=dp, rlc_cw=dp * 0.5, sigma_qcw=rsd, avg_cf=frac,
                        p_layer=st["p_layer"], xcw=xcw, ref_c_sub=c_sub, ref_n_subcol_cld=ncld)


TWO_STREAM = {
    # name: (n_profile, n_layer, isolir, l_scale_solar, l_ir_source_quad, seed, omc_range)
    "sw_scaled": (8, 40, 1, 1, 1, 20261208, (-8.0, -1.0)),
    "sw_plain": (8, 40, 1, 0, 1, 20261209, (-3.0, 0.0)),
    "lw_quad": (8, 40, 2, 1, 1, 20261210, (-3.0, 0.0)),
    "lw_linear": (8, 40, 2, 1, 0, 20261211, (-3.0, 0.0)),
}


def run_two_stream(name, P, L, isolir, lss, lq, seed, omc, tmp):
    tsc = synth.make_two_stream_case(P, L, isolir, seed, l_scale_solar=lss, l_ir_source_quad=lq,
                                     omc_range=omc)
    fin, fout = os.path.join(tmp, "ts.bin"), os.path.join(tmp, "ts.out")
    synth.write_two_stream_case(tsc, fin)
    subprocess.run([os.path.join(REF, "soc_two_stream_case"), fin, fout], check=True)
    a = np.fromfile(fout)
    fdir = a[:P * (L + 1)].reshape(L + 1, P)
    ftot = a[P * (L + 1):].reshape(2 * L + 2, P)
    t = synth.two_stream_term(tsc, 0)
    np.savez_compressed(os.path.join(HERE, "two_stream_%s.npz" % name),
                        **{k: np.asarray(v) for k, v in t.items()},
                        ref_flux_direct=fdir, ref_flux_total=ftot)


def main():
    with tempfile.TemporaryDirectory() as tmp:
        for name, (P, L, src, cloudy, seed) in CASES.items():
            run_case(name, P, L, src, cloudy, seed, tmp)
        for ov in (0, 1, 2, 3):
            run_cloud_gen(ov, tmp)
        for name, args in TWO_STREAM.items():
            run_two_stream(name, *args, tmp)
    ver = subprocess.run(["/opt/rocm/bin/amdflang", "--version"], capture_output=True,
                         text=True).stdout.splitlines()[0]
    json.dump({"compiler": ver, "flags": "-O2 -fPIC -ffp-contract=off",
               "cases": {k: list(v) for k, v in CASES.items()},
               "two_stream_cases": {k: [list(x) if isinstance(x, tuple) else x for x in v]
                                    for k, v in TWO_STREAM.items()},
               "spectral_files": ["socrates_amd/data/sp_lw_synth", "socrates_amd/data/sp_sw_synth",
                                  "socrates_amd/data/mcica_synth"]},
              open(os.path.join(HERE, "meta.json"), "w"), indent=1)
    print("golden vectors written")


if __name__ == "__main__":
    main()
