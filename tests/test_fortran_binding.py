"""The Fortran binding (socrates_amd/fortran/socrates_runes_gpu.f90): a host
program written against the reference's module and keyword names
(set_spectrum, set_mcica, runes, StrDiag) runs on the device and matches the
oracle to the FP64 tolerances of tests/test_gpu_parity.py; its mixed form
stays within the mixed bounds. Also: the grey-albedo option of set_bound."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle
from socrates_amd import runes as R
from socrates_amd import synth
from socrates_amd.case_io import read_output
from test_gpu_parity import MCICA, NBAND, SPEC, check_fp64, check_mixed

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "socrates_amd", "lib", "runes_gpu_case")


def test_fortran_binding_built():
    assert os.path.exists(EXE), "build the Fortran binding (python socrates_amd/build.py)"


@pytest.mark.gpu
@pytest.mark.parametrize("src,cloudy", [("lw", True), ("sw", True), ("sw", False)])
def test_fortran_binding_matches_oracle(src, cloudy):
    case = synth.make_case(12, 50, src, NBAND[src], seed=900 + cloudy, cloudy=cloudy)
    ref = oracle.runes(case, SPEC[src], MCICA if cloudy else None)
    with tempfile.TemporaryDirectory() as tmp:
        cf = os.path.join(tmp, "case.bin")
        case.write(cf)
        for prec in ("fp64", "mixed"):
            of = os.path.join(tmp, "out_%s.bin" % prec)
            r = subprocess.run([EXE, SPEC["sw"], SPEC["lw"], MCICA if cloudy else "-", cf, of, prec],
                               capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stdout + r.stderr
            out = read_output(of, case.n_profile, case.n_layer)
            if prec == "fp64":
                check_fp64(out, ref, "fortran %s" % src, src, case.layer_heat_capacity)
            else:
                check_mixed(out, ref, "fortran mixed %s" % src)


@pytest.mark.gpu
@pytest.mark.parametrize("src", ["lw", "sw"])
def test_fortran_binding_ga9_with_further_gases(src):
    """The binding on the GA9 files: the further gases reach the library
    through runes' *_mix_ratio keywords."""
    from ga9_cases import MCICA_GA9, SPEC_GA9
    case = synth.make_case(10, 60, src, NBAND[src], seed=950 + (src == "sw"), cloudy=True)
    case.wm_mix_ratio = np.array([4.7e-7, 9.6e-7, 0.23, 2e-10, 8e-10, 1.1e-9, 2.4e-9, 3.3e-10, 9e-10, 4.5e-10])
    ref = oracle.runes(case, SPEC_GA9[src], MCICA_GA9)
    with tempfile.TemporaryDirectory() as tmp:
        cf, of = os.path.join(tmp, "case.bin"), os.path.join(tmp, "out.bin")
        case.write(cf)
        r = subprocess.run([EXE, SPEC_GA9["sw"], SPEC_GA9["lw"], MCICA_GA9, cf, of, "fp64"],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        out = read_output(of, case.n_profile, case.n_layer)
        check_fp64(out, ref, "fortran ga9 %s" % src, src, case.layer_heat_capacity)


@pytest.mark.gpu
@pytest.mark.parametrize("src", ["lw", "sw"])
def test_fortran_binding_l_invert(src):
    """Bottom-up profiles with l_invert = .true. (socrates_set_atm.F90:
    218-225, socrates_set_diag.F90:113-151) and profile-last fields with
    l_profile_last = .true. give the plain call's results, flipped or
    transposed: the example program rearranges its inputs, calls runes with
    the flag and rearranges the outputs back, to the same bits."""
    case = synth.make_case(9, 40, src, NBAND[src], seed=970 + (src == "sw"), cloudy=True)
    with tempfile.TemporaryDirectory() as tmp:
        cf = os.path.join(tmp, "case.bin")
        case.write(cf)
        outs = []
        for mode in ("plain", "invert", "last"):
            of = os.path.join(tmp, "out_%s.bin" % mode)
            r = subprocess.run([EXE, SPEC["sw"], SPEC["lw"], MCICA, cf, of, "fp64", mode],
                               capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stdout + r.stderr
            outs.append(read_output(of, case.n_profile, case.n_layer))
        for k in ("flux_up", "flux_down", "flux_direct", "heating_rate"):
            assert np.array_equal(outs[0][k], outs[1][k]), k
            assert np.array_equal(outs[0][k], outs[2][k]), k


@pytest.mark.gpu
def test_grey_albedo_is_every_band():
    R.register_default_spectra()
    case = synth.make_case(6, 40, "sw", NBAND["sw"], seed=77, cloudy=False)
    grey = 0.23
    case.albedo_diff[:] = grey
    case.albedo_dir[:] = grey
    ref = oracle.runes(case, SPEC["sw"], None)
    out = R.runes(case.n_profile, case.n_layer, "sw", 1, case.p_layer, case.t_layer, case.mass,
                  density=case.density, layer_heat_capacity=case.layer_heat_capacity,
                  h2o=case.h2o, o3=case.o3, co2_mix_ratio=case.co2_mix_ratio,
                  p_level=case.p_level, t_level=case.t_level, t_ground=case.t_ground,
                  cos_zenith_angle=case.cos_zen, solar_irrad=case.solar_irrad,
                  l_grey_albedo=True, grey_albedo=grey, l_rayleigh=case.l_rayleigh)
    check_fp64(out, ref, "grey albedo", "sw", case.layer_heat_capacity)


@pytest.mark.gpu
def test_python_l_invert_matches_top_down():
    """runes(..., l_invert=True) on bottom-up arrays gives the top-down
    call's outputs, flipped, to the bit."""
    from socrates_amd import runes as R
    R.register_default_spectra()
    case = synth.make_case(6, 35, "lw", NBAND["lw"], seed=975, cloudy=True)
    name, kw = R.case_call(case, "lw")
    P, L, src = kw.pop("n_profile"), kw.pop("n_layer"), kw.pop("i_source")
    top = R.runes(P, L, name, src, precision="fp64", **kw)
    flip = {k: (np.ascontiguousarray(v[::-1]) if isinstance(v, np.ndarray) and v.ndim == 2 and
                k not in ("albedo_diff", "albedo_dir") else v) for k, v in kw.items()}
    bot = R.runes(P, L, name, src, precision="fp64", l_invert=True, **flip)
    for k in ("flux_up", "flux_down", "heating_rate"):
        assert np.array_equal(np.asarray(bot[k])[::-1], np.asarray(top[k])), k
