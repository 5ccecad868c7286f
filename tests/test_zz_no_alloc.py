"""The allocation-free entry point (socgpu_runes_noalloc) against the
ordinary one (socgpu_runes) through the C ABI: the same outputs, bit for
bit, on two calls in a row (the second smaller, so stale workspace from the
first would show), both forms, solar and infra-red all-sky McICA calls with
inputs and outputs in HBM; a call above the fixed workspace and host arrays
are refused."""
import numpy as np
import pytest

from socrates_amd import runes as R
from socrates_amd import synth

pytestmark = pytest.mark.gpu

NBAND = {"lw": 9, "sw": 6}
FIELDS = ("p_layer", "t_layer", "mass", "density", "layer_heat_capacity", "h2o", "o3", "p_level",
          "t_level", "t_ground", "solar_irrad", "albedo_diff", "albedo_dir", "cloud_frac", "liq_mmr",
          "liq_dim", "rand_seed")


@pytest.fixture(scope="module", autouse=True)
def spectra():
    R.register_default_spectra()


def _device(case):
    import torch
    dev = torch.device("cuda:0")
    arrs = {k: torch.from_numpy(np.ascontiguousarray(getattr(case, k))).to(dev) for k in FIELDS}
    arrs["cos_zenith_angle"] = torch.from_numpy(np.ascontiguousarray(case.cos_zen)).to(dev)
    return arrs


def _outs(case):
    import torch
    L, P = case.n_layer, case.n_profile
    return {k: torch.full(((L + 1) if k != "heating_rate" else L, P), -1.0, dtype=torch.float64,
                          device="cuda:0")
            for k in ("flux_up", "flux_down", "flux_direct", "heating_rate")}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("src", ["sw", "lw"])
def test_noalloc_matches_runes(precision, src):
    import torch
    for k, (P, L) in enumerate(((300, 90), (37, 61))):
        case = synth.make_case(P, L, src, NBAND[src], seed=4100 + k + (src == "sw"), cloudy=True)
        arrs = _device(case)
        name, kw = R.case_call(case, src, device_arrays=arrs, out=_outs(case))
        ref = R.runes(kw.pop("n_profile"), kw.pop("n_layer"), name, kw.pop("i_source"),
                      precision=precision, **kw)
        name, kw = R.case_call(case, src, device_arrays=arrs, out=_outs(case))
        got = R.runes_noalloc(kw.pop("n_profile"), kw.pop("n_layer"), name, kw.pop("i_source"),
                              precision=precision, **kw)
        torch.cuda.synchronize()
        for f in ("flux_up", "flux_down", "flux_direct", "heating_rate"):
            assert torch.equal(got[f], ref[f]), (precision, src, P, L, f)


def test_noalloc_refuses_oversized_call_and_host_arrays():
    big = synth.make_case(4096, 137, "lw", 9, seed=3, cloudy=True)
    arrs = _device(big)
    name, kw = R.case_call(big, "lw", device_arrays=arrs, out=_outs(big))
    with pytest.raises(R.SocgpuError, match="fixed workspace"):
        R.runes_noalloc(kw.pop("n_profile"), kw.pop("n_layer"), name, kw.pop("i_source"), **kw)
    # host arrays: refused by the library itself
    import ctypes
    small = synth.make_case(4, 30, "lw", 9, seed=4, cloudy=True)
    name, kw = R.case_call(small, "lw")
    a, d, out, keep = R._prepare(kw.pop("n_profile"), kw.pop("n_layer"), name, kw.pop("i_source"), **kw)
    rc = R.lib().socgpu_runes_noalloc(name.encode(), ctypes.byref(a), ctypes.byref(d), 0)
    assert rc == -1 and b"device inputs" in R.lib().socgpu_last_error()
    # the entry point still works after a refusal
    c = synth.make_case(8, 40, "sw", 6, seed=5, cloudy=True)
    a = _device(c)
    name, kw = R.case_call(c, "sw", device_arrays=a, out=_outs(c))
    got = R.runes_noalloc(kw.pop("n_profile"), kw.pop("n_layer"), name, kw.pop("i_source"), **kw)
    name, kw = R.case_call(c, "sw", device_arrays=a, out=_outs(c))
    ref = R.runes(kw.pop("n_profile"), kw.pop("n_layer"), name, kw.pop("i_source"), **kw)
    import torch
    assert torch.equal(got["flux_down"], ref["flux_down"])
