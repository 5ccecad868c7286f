#!/usr/bin/env python3
"""Build the reference SOCRATES radiation core (Fortran, FP64) into oracle/_ref/.

Test infrastructure only: the library built here is the checker that pins the
C restatement in oracle/ and the CPU baseline in bench.py. Nothing in the
product links it.

The recipe compiles the reference sources where they lie
(src/modules_core, src/radiance_core, src/interface_core under the reference
tree) with amdflang, strict FP (-O2, no fast-math), in module-dependency order
computed from their USE statements, and links:
  oracle/_ref/libsocref.so   - radiance core + runes interface (shared)
  oracle/_ref/<harness>      - the harness programs in oracle/harness/*.f90
It never runs the reference's own build system and never copies its sources.

Usage: python3 oracle/build_ref.py [--ref /root/reference] [-j 8]
"""
import argparse
import concurrent.futures as cf
import glob
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
DIRS = ["modules_core", "radiance_core", "interface_core"]
FC = os.environ.get("SOC_FC", "/opt/rocm/bin/amdflang")
FFLAGS = ["-O2", "-fPIC", "-ffp-contract=off"]
# reference programs built as-is: the runes test driver behind examples/runes
REF_PROGRAMS = ["aux/runes_driver.f90"]
# host programs of the Fortran binding that also build against the
# reference, so that both read the same caller arrays
HOST_PROGRAMS = [os.path.join(os.path.dirname(HERE), "socrates_amd", "fortran", "runes_lfric_case.f90")]

_mod_re = re.compile(r"^\s*module\s+(\w+)\s*(?:!.*)?$", re.I | re.M)
_use_re = re.compile(r"^\s*use\s*(?:,\s*intrinsic\s*::)?\s*(\w+)", re.I | re.M)


def scan(files):
    defs, uses = {}, {}
    for f in files:
        txt = open(f, errors="replace").read()
        for m in _mod_re.finditer(txt):
            if m.group(1).lower() != "procedure":
                defs[m.group(1).lower()] = f
        uses[f] = {m.group(1).lower() for m in _use_re.finditer(txt)}
    return defs, uses


def levels(files, defs, uses):
    """Group files into compile levels: each level only USEs earlier levels."""
    dep = {f: {defs[u] for u in uses[f] if u in defs and defs[u] != f} for f in files}
    done, out = set(), []
    while len(done) < len(files):
        lvl = [f for f in files if f not in done and dep[f] <= done]
        if not lvl:
            raise RuntimeError("cyclic module dependencies: %s" % sorted(set(files) - done))
        out.append(sorted(lvl))
        done.update(lvl)
    return out


def compile_one(src, moddir, objdir):
    obj = os.path.join(objdir, os.path.basename(src).rsplit(".", 1)[0] + ".o")
    if os.path.exists(obj) and os.path.getmtime(obj) > os.path.getmtime(src):
        return obj
    cmd = [FC] + FFLAGS + ["-module-dir", moddir, "-I", moddir, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    return obj


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    src_root = os.path.join(a.ref, "src")
    if not os.path.isdir(src_root):
        print("reference tree not found at %s; nothing to build" % src_root)
        return 0
    moddir = os.path.join(OUT, "mod")
    objdir = os.path.join(OUT, "obj")
    os.makedirs(moddir, exist_ok=True)
    os.makedirs(objdir, exist_ok=True)
    files = []
    for d in DIRS:
        files += sorted(glob.glob(os.path.join(src_root, d, "*.[Ff]90")))
        files += sorted(glob.glob(os.path.join(src_root, d, "*.f")))
    defs, uses = scan(files)
    objs = []
    with cf.ThreadPoolExecutor(a.j) as ex:
        for lvl in levels(files, defs, uses):
            objs += list(ex.map(lambda s: compile_one(s, moddir, objdir), lvl))
    lib = os.path.join(OUT, "libsocref.so")
    subprocess.run([FC, "-shared", "-o", lib] + objs, check=True)
    # the reference's own driver program (compiled where it lies) and our
    # harness programs, linked against the reference library
    progs = [os.path.join(src_root, p) for p in REF_PROGRAMS]
    progs += sorted(glob.glob(os.path.join(HERE, "harness", "*.f90"))) + HOST_PROGRAMS
    for h in progs:
        exe = os.path.join(OUT, os.path.basename(h)[:-4])
        cmd = [FC] + FFLAGS + ["-module-dir", objdir, "-I", moddir, h, "-o", exe,
                               "-L" + OUT, "-lsocref", "-Wl,-rpath," + OUT]
        subprocess.run(cmd, check=True)
    print("built %s (%d objects)" % (lib, len(objs)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
