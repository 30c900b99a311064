"""Build libmw3.so (HIP, gfx950) in-tree.

Every source is compiled with hipcc for gfx950 only; packed FP32 instructions
are disabled (-packed-fp32-ops) for all device code.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmw3.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["mw3_setup.cpp", "mw3_kernels.hip", "mw3_api.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _compile(src, obj, extra):
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC] + FLAGS + extra + lang + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s%s" % (src, r.stdout, r.stderr))
    return obj


def build(verbose=False, extra_flags=()):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        futs = []
        for s in SOURCES:
            src = os.path.join(CSRC, s)
            obj = os.path.join(objdir, os.path.splitext(s)[0] + ".o")
            futs.append(ex.submit(_compile, src, obj, list(extra_flags)))
        for f in futs:
            objs.append(f.result())
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FDRIVE = os.path.join(FDIR, "mw3_fdrive")


def build_fortran(verbose=False):
    """Compile the ISO_C_BINDING module mw3_capi and the Fortran host program
    mw3_fdrive against libmw3.so (amdflang)."""
    objdir = os.path.join(HERE, "_build", "fortran")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in ("mw3_capi.f90", "mw3_fdrive.f90"):
        obj = os.path.join(objdir, src.replace(".f90", ".o"))
        cmd = [FLANG, "-O2", "-fPIC", "-module-dir", objdir, "-c", os.path.join(FDIR, src),
               "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("amdflang failed for %s:\n%s%s" % (src, r.stdout, r.stderr))
        objs.append(obj)
    tmp = FDRIVE + ".tmp"
    cmd = [FLANG, "-o", tmp] + objs + ["-L" + HERE, "-lmw3", "-L/opt/rocm/lib",
                                       "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("Fortran link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, FDRIVE)
    if verbose:
        print("built", FDRIVE)
    return FDRIVE


if __name__ == "__main__":
    build(verbose=True)
    build_fortran(verbose=True)
    sys.exit(0)
