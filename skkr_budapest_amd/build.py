"""Build steps of the package: the HIP library (gfx950) and the Fortran host.

The library is compiled in-tree (skkr_budapest_amd/libskkr_gpu.so) so that it
travels with the source tree; hipcc cross-compiles without a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")

HIPCC = os.environ.get("HIPCC", "hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Wno-unused-result", "-Wno-unused-value",
             # packed FP32 instructions stay off (host pass warns, harmless)
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build_hip(verbose=False):
    out = os.path.join(HERE, "libskkr_gpu.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "skkr_gpu.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_mlp(verbose=False):
    out = os.path.join(HERE, "libskkr_mlp.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "mlp.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_fortran(verbose=False):
    """Fortran host: ISO_C_BINDING module + cpacoord-style k-loop driver."""
    fc = shutil.which("flang") or shutil.which("amdflang")
    if fc is None or not os.path.isdir(FSRC):
        return None
    bdir = os.path.join(FSRC, "_build")
    os.makedirs(bdir, exist_ok=True)
    srcs = [os.path.join(FSRC, f) for f in ("skkr_gpu_mod.f90", "kloop_gpu.f90",
                                             "skkr_kloop_main.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    exe = os.path.join(FSRC, "skkr_kloop")
    cmd = [fc, "-O2", "-J", bdir, *srcs, "-o", exe,
           "-L" + HERE, "-lskkr_gpu", "-Wl,-rpath," + HERE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("flang failed:\n" + r.stdout + r.stderr)
    if verbose:
        print("built", exe)
    # the structure-constant producer, tau2d, and the reference-shaped dugo /
    # tau2d (skkr_dropin.f90) called from Fortran
    for extra, src, name in (((), "skkr_gstore_main.f90", "skkr_gstore"),
                             ((), "skkr_tau2d_main.f90", "skkr_tau2d"),
                             (("skkr_dropin.f90",), "skkr_g2d_main.f90", "skkr_g2d")):
        gsrc = os.path.join(FSRC, src)
        if not os.path.exists(gsrc):
            continue
        gexe = os.path.join(FSRC, name)
        cmd = [fc, "-O2", "-J", bdir, srcs[0], *(os.path.join(FSRC, e) for e in extra), gsrc, "-o", gexe,
               "-L" + HERE, "-lskkr_gpu", "-Wl,-rpath," + HERE]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("flang failed:\n" + r.stdout + r.stderr)
        if verbose:
            print("built", gexe)
    return exe


def build_all(verbose=False):
    build_hip(verbose)
    build_mlp(verbose)
    build_fortran(verbose)
