"""Build recipe for libexh.so (hipcc, gfx950) -- compiled in-tree.

    python -m exciting_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libexh.so")
SOURCES = ["api.hip", "setup_kernels.hip", "match_kernel.hip", "ozaki_kernels.hip", "batch.hip",
           "interstitial.hip", "hmlint.hip", "dense_kernels.hip"]
# hipFFT for the 3-D transforms of the interstitial producers
LINK = ["-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]
ARCH = os.environ.get("EXH_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            # packed FP32 VALU ops stay off in all device code
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            # keep MFMA accumulators in the unified VGPR file: the default
            # heuristic copies every loop-carried accumulator AGPR<->VGPR per
            # iteration of the K loop
            "-mllvm", "-amdgpu-mfma-vgpr-form=1",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
            "-I", os.path.join(HERE, "..", "include")]


def build(verbose: bool = False) -> str:
    objs = []
    cc = hipcc()
    procs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [cc, *flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode(errors="replace"))
            raise RuntimeError(f"hipcc failed on {src}")
        msg = out.decode(errors="replace")
        if verbose and msg.strip():
            sys.stderr.write(msg)
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, *LINK]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "exh_driver")
SHIM_DRIVER = os.path.join(FORTRAN, "exh_shim_driver")
# the binding module, the caller's matrix types (standalone build only), the
# integration's state module and the reference-name shims
MODULES = ("exh_lapw.f90", "fvsystem.f90", "exh_state.f90", "exh_shims.f90")


def flang() -> str:
    for cand in (os.environ.get("FLANG"), "/opt/rocm/lib/llvm/bin/flang", shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    return ""


def build_fortran(verbose: bool = False) -> str:
    """The ISO_C_BINDING module, the drop-in shims (match,
    hamiltonandoverlapsetup, overlapsetup under the reference's names) and
    the two driver programs, linked to libexh.so.  Skipped (returns "") when
    no Fortran compiler is installed."""
    fc = flang()
    if not fc:
        return ""
    bdir = os.path.join(FORTRAN, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = {}
    for src in MODULES + ("exh_driver.f90", "exh_shim_driver.f90"):
        obj = os.path.join(bdir, src.replace(".f90", ".o"))
        cmd = [fc, "-O2", "-c", os.path.join(FORTRAN, src), "-module-dir", bdir, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs[src] = obj
    for prog, main, deps in ((DRIVER, "exh_driver.f90", ("exh_lapw.f90",)),
                             (SHIM_DRIVER, "exh_shim_driver.f90", MODULES)):
        tmp = prog + ".tmp"
        cmd = [fc, "-o", tmp, objs[main], *[objs[d] for d in deps], "-L", HERE, "-lexh",
               "-Wl,-rpath,$ORIGIN/.."]
        subprocess.run(cmd, check=True)
        os.replace(tmp, prog)
    shutil.rmtree(bdir, ignore_errors=True)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True) or "no Fortran compiler: driver skipped")
