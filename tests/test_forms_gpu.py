"""The further forms of runes on the device, through the C ABI:

* CO2 and further-gas fields, the surface emission per band and the
  clear-sky diagnostics against the reference's golden vectors
  (tests/golden/forms_ga9_*.npz) to the FP64 tolerances of
  tests/test_gpu_parity.py, and within the mixed bounds;
* clear-sky fluxes bit-identical to the same call with cloud off, from the
  single call, the packed batch, the persistent batch and the serving loop,
  with host and device outputs;
* H2O and O3 as well-mixed ratios: bit-identical to constant fields;
* the Fortran binding in the host-model calling form (runes_lfric_case
  built against the binding) against the same program built against the
  reference (tests/golden/lfric_*.npz), gas selection included.
"""
import dataclasses
import glob
import os
import subprocess
import tempfile

import numpy as np
import pytest

from socrates_amd import runes as R
from socrates_amd import synth
from socrates_amd.case_io import case_from_arrays
from test_forms import CLEAR, OUTS, lfric_parse
from test_gpu_parity import MCICA, NBAND, SPEC, check_fp64, check_mixed, oracle_spread

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
GA9 = os.path.join(ROOT, "socrates_amd", "data", "ga9")
SPEC_GA9 = {"sw": os.path.join(GA9, "sp_sw_ga9"), "lw": os.path.join(GA9, "sp_lw_ga9")}
MCICA_GA9 = os.path.join(GA9, "mcica_data")
LFRIC_EXE = os.path.join(ROOT, "socrates_amd", "lib", "runes_lfric_case")


@pytest.fixture(scope="module", autouse=True)
def spectra():
    R.register_default_spectra()
    R.set_spectrum("ga9_sw", SPEC_GA9["sw"])
    R.set_spectrum("ga9_lw", SPEC_GA9["lw"])
    R.set_mcica(MCICA_GA9, sw_spectrum_name="ga9_sw", lw_spectrum_name="ga9_lw")


def _clear_ref(d):
    return {k: d["ref_" + k + "_clear"] for k in ("flux_up", "flux_down", "flux_direct")}


@pytest.mark.parametrize("src", ["sw", "lw"])
def test_forms_match_reference(src):
    d = np.load(os.path.join(GOLD, "forms_ga9_%s.npz" % src))
    case = case_from_arrays(d)
    ref = {k: d["ref_" + k] for k in OUTS}
    name, kw = R.case_call(case, "ga9_" + src)
    kw = dict(kw, l_clear=True)
    P, L, s = kw.pop("n_profile"), kw.pop("n_layer"), kw.pop("i_source")
    out = R.runes(P, L, name, s, precision="fp64", **kw)
    check_fp64(out, ref, "forms %s" % src, src, case.layer_heat_capacity)
    cl = {k: out[k + "_clear"] for k in ("flux_up", "flux_down", "flux_direct")}
    cref = _clear_ref(d)
    for k in cl:
        r = cref[k]
        tol = np.maximum(1e-10 * np.abs(r), 1e-9)
        assert (np.abs(cl[k] - r) <= tol).all(), (k, np.abs(cl[k] - r).max())
    mixed = R.runes(P, L, name, s, precision="mixed", **kw)
    check_mixed(mixed, ref, "forms mixed %s" % src)


def _cloud_off(case):
    return dataclasses.replace(case, l_cloud=0, i_inhom=0)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", ["mcica", "triple"])
def test_clear_sky_equals_cloud_off_call(prec, kind):
    cases = [synth.make_case(9, 40, src, NBAND[src], seed=1400 + i, cloudy=True, overlap=kind)
             for i, src in enumerate(("sw", "lw"))]
    for case, src in zip(cases, ("sw", "lw")):
        name, kw = R.case_call(case, src)
        kw = dict(kw, l_clear=True)
        P, L, s = kw.pop("n_profile"), kw.pop("n_layer"), kw.pop("i_source")
        out = R.runes(P, L, name, s, precision=prec, **kw)
        off = R.runes_case(_cloud_off(case), src, precision=prec)
        for k in ("flux_up", "flux_down", "flux_direct"):
            assert np.array_equal(out[k + "_clear"], off[k]), (src, k)
        # the cloudy outputs are those of the call without the diagnostics
        plain = R.runes_case(case, src, precision=prec)
        for k in OUTS:
            assert np.array_equal(out[k], plain[k]), (src, k)


def test_clear_sky_through_batches_and_serving():
    from socrates_amd.serve import Server
    calls, offs = [], []
    for i, (src, kind) in enumerate((("sw", "mcica"), ("lw", "triple"), ("lw", "mcica"), ("sw", "clear"))):
        case = synth.make_case(5 + i, 30 + 3 * i, src, NBAND[src], seed=1500 + i, cloudy=kind != "clear",
                               overlap=kind if kind != "clear" else "mcica")
        name, kw = R.case_call(case, src)
        calls.append((name, dict(kw, l_clear=(i != 2))))
        offs.append(R.runes_case(_cloud_off(case), src, precision="fp64"))
    for run in ("batch", "persistent", "serve"):
        if run == "batch":
            outs = R.runes_batch(calls)
        elif run == "persistent":
            b = R.Batch()
            b.run(calls)
            outs = b.run(calls)
            b.close()
        else:
            srv = Server(precision="fp64", hbm_budget_gb=4)
            ts = [srv.submit(n, kw) for n, kw in calls]
            outs = [srv.wait(t) for t in ts]
            srv.close()
        for i, (o, off) in enumerate(zip(outs, offs)):
            if i == 2:
                assert "flux_up_clear" not in o
                continue
            for k in ("flux_up", "flux_down", "flux_direct"):
                assert np.array_equal(o[k + "_clear"], off[k]), (run, i, k)


def test_clear_sky_device_outputs():
    import torch
    case = synth.make_case(8, 35, "sw", 6, seed=1601, cloudy=True)
    case.fill_defaults()
    fields = ("p_layer", "t_layer", "mass", "density", "layer_heat_capacity", "h2o", "o3", "p_level",
              "t_level", "t_ground", "solar_irrad", "albedo_diff", "albedo_dir", "cloud_frac", "liq_mmr",
              "liq_dim", "rand_seed")
    dev = {k: torch.from_numpy(np.ascontiguousarray(getattr(case, k))).cuda() for k in fields}
    dev["cos_zenith_angle"] = torch.from_numpy(np.ascontiguousarray(case.cos_zen)).cuda()
    name, kw = R.case_call(case, "sw", device_arrays=dev)
    kw = dict(kw, l_clear=True)
    P, L, s = kw.pop("n_profile"), kw.pop("n_layer"), kw.pop("i_source")
    out = R.runes(P, L, name, s, **kw)
    off = R.runes_case(_cloud_off(case), "sw")
    for k in ("flux_up", "flux_down", "flux_direct"):
        assert np.array_equal(out[k + "_clear"].cpu().numpy(), off[k]), k


def test_well_mixed_h2o_o3_equal_constant_fields():
    case = synth.make_case(6, 30, "lw", 9, seed=1700, cloudy=False)
    name, kw = R.case_call(case, "lw")
    P, L, s = kw.pop("n_profile"), kw.pop("n_layer"), kw.pop("i_source")
    field = R.runes(P, L, name, s, **dict(kw, h2o=np.full((L, P), 3.0e-3), o3=np.full((L, P), 2.0e-6)))
    well = R.runes(P, L, name, s, **dict(kw, h2o=None, o3=None, h2o_mix_ratio=3.0e-3, o3_mix_ratio=2.0e-6))
    for k in OUTS:
        assert np.array_equal(field[k], well[k]), k
    dry = R.runes(P, L, name, s, **dict(kw, h2o=None, o3=None))
    assert np.abs(dry["flux_up"] - well["flux_up"]).max() > 1e-3


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "lfric_*.npz"))))
def test_fortran_host_model_form_matches_reference(path):
    d = np.load(path)
    case = case_from_arrays(d)
    P, L = case.n_profile, case.n_layer
    src = "sw" if case.i_source == 1 else "lw"
    want = lfric_parse(d["lfric_raw"], P, L)
    with tempfile.TemporaryDirectory() as t:
        cf, of = os.path.join(t, "case.bin"), os.path.join(t, "out.bin")
        case.write(cf)
        mc = MCICA if case.l_cloud and case.i_cloud_repr == 1 else "-"
        args = [LFRIC_EXE, SPEC["sw"], SPEC["lw"], mc, cf, of]
        if int(d["drop"]):
            args.append(str(int(d["drop"])))
        print("case", os.path.basename(path), flush=True)
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        got = lfric_parse(np.fromfile(of), P, L)
    # with modal aerosols in optically thin infra-red layers the reference's
    # own fluxes move by up to ~2e-7 W m-2 under a one-ulp change of exp:
    # that measured movement (four jittered oracle runs, twice) is added
    spread = (oracle_spread(case, src, case.l_cloud and case.i_cloud_repr == 1, runs=4)
              if case.n_aer_mode > 0 else None)
    check_fp64(got, {k: want[k] for k in OUTS}, os.path.basename(path), src, case.layer_heat_capacity,
               spread=spread, factor=2.0)
    for k in CLEAR + ("up_surf", "down_surf", "direct_surf", "up_toa", "down_toa", "direct_toa"):
        r = want[k]
        tol = np.maximum(1e-10 * np.abs(r), 1e-9)
        if spread is not None:
            tol = tol + 2.0 * spread["flux_up"].max()
        assert (np.abs(got[k] - r) <= tol).all(), k
    for k in ("hr_above", "hr_below", "unused_columns"):
        assert np.array_equal(got[k], want[k]), k
