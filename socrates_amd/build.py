"""Build the HIP library socrates_amd/lib/libsocgpu.so for gfx950 with hipcc,
and the Fortran binding (socrates_amd/fortran) with amdflang: its module
files and object in socrates_amd/lib/fortran, and the example host programs
socrates_amd/lib/runes_gpu_case and runes_lfric_case linked against
libsocgpu.so.

All HIP code is compiled with packed FP32 instructions disabled
(-target-feature -packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "lib", "libsocgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = ["socgpu.hip", "spectrum.cpp"]
FC = os.environ.get("SOCGPU_FC", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FOUT = os.path.join(HERE, "lib", "fortran")


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + [os.path.join(SRC, s) for s in SOURCES]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libsocgpu.so")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    build_fortran(verbose)
    return OUT


def build_fortran(verbose=False):
    """The Fortran binding and its example program (skipped without amdflang)."""
    if not os.path.exists(FC):
        if verbose:
            print("amdflang not found: Fortran binding not built")
        return None
    os.makedirs(FOUT, exist_ok=True)
    obj = os.path.join(FOUT, "socrates_runes_gpu.o")
    steps = [
        [FC, "-O2", "-fPIC", "-c", os.path.join(FDIR, "socrates_runes_gpu.f90"),
         "-module-dir", FOUT, "-o", obj],
        [FC, "-O2", "-I", FOUT, os.path.join(FDIR, "runes_gpu_case.f90"), obj,
         "-L" + os.path.dirname(OUT), "-lsocgpu", "-Wl,-rpath,$ORIGIN",
         "-o", os.path.join(HERE, "lib", "runes_gpu_case")],
        [FC, "-O2", "-I", FOUT, os.path.join(FDIR, "runes_lfric_case.f90"), obj,
         "-L" + os.path.dirname(OUT), "-lsocgpu", "-Wl,-rpath,$ORIGIN",
         "-o", os.path.join(HERE, "lib", "runes_lfric_case")],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=FOUT)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("amdflang failed: " + " ".join(cmd[:4]))
    if verbose:
        print("built", os.path.join(HERE, "lib", "runes_gpu_case"))
    return obj


if __name__ == "__main__":
    build(verbose=True)
