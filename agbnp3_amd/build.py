"""Builds every native library of the package in-tree.

  lib/libagbnp3.so        the engine (HIP for gfx950 + host C-ABI), float_i = double
  lib/libagbnp3_sp.so     the same engine behind the single-precision interface
                          (agbnp3.h built with SINGLE_PREC_INTERFACE: float_i = float)
  lib/libnblist.so        NeighList container (C)
  lib/libagbnp3_synth.so  seeded synthetic solutes (C)
and, as test infrastructure, oracle/liboracle.so plus the reference build
(oracle/_ref/) when the reference sources are present.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

from . import _paths

HIP_SOURCES = ["agbnp3_host.hip", "k_setup.hip", "k_tree.hip", "k_pairs.hip", "k_epi.hip", "k_hb.hip", "k_surrogate.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AGBNP3_ARCH", "gfx950")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions (they have given silent wrong results on these GPUs)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
]
# Translation units whose kernels must reproduce the reference's FP32
# distance tests exactly: contraction is allowed only where the source permits
# it, so `#pragma clang fp contract(off)` is honoured (the default "fast" mode
# fuses across statements and ignores the pragma).
HIP_FILE_FLAGS = {
    "k_setup.hip": ["-ffp-contract=on"],
    "k_hb.hip": ["-ffp-contract=on"],
}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    elif p.stderr.strip() and verbose:
        sys.stderr.write(p.stderr)
    return p


def build_c(verbose=True):
    os.makedirs(_paths.LIB_DIR, exist_ok=True)
    csrc, inc = _paths.CSRC_DIR, _paths.INCLUDE_DIR
    _run(["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-o", _paths.LIBSYNTH, os.path.join(csrc, "synth.c"), "-lm"],
         verbose)
    _run(["gcc", "-O2", "-fPIC", "-shared", "-Wall", f"-I{inc}", "-o", _paths.LIBNBLIST,
          os.path.join(csrc, "nblist.c")], verbose)


def build_hip(verbose=True, jobs=6):
    objdir = os.path.join(_paths.LIB_DIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    inc = [f"-I{_paths.INCLUDE_DIR}", f"-I{_paths.CSRC_DIR}"]
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC] + HIP_FLAGS + HIP_FILE_FLAGS.get(src, []) + inc +
                    ["-c", os.path.join(_paths.CSRC_DIR, src), "-o", obj])
    # the host C-ABI once more with float_i = float
    host_sp = os.path.join(objdir, "agbnp3_host_sp.o")
    cmds.append([HIPCC] + HIP_FLAGS + ["-DSINGLE_PREC_INTERFACE"] + inc +
                ["-c", os.path.join(_paths.CSRC_DIR, "agbnp3_host.hip"), "-o", host_sp])
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    kernels = objs[1:]
    for lib, host in ((_paths.LIBAGBNP3, objs[0]), (_paths.LIBAGBNP3_SP, host_sp)):
        tmp = lib + ".tmp"
        _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, host] + kernels +
             [f"-L{_paths.LIB_DIR}", "-lnblist", "-Wl,-rpath,$ORIGIN"], verbose)
        os.replace(tmp, lib)


def build_oracle(verbose=True):
    odir = os.path.join(_paths.REPO_DIR, "oracle")
    _run(["make", "-C", odir, "oracle"], verbose)
    ref = os.environ.get("AGBNP3_REF_DIR", "/root/reference")
    if os.path.isdir(ref) and shutil.which("gcc"):
        _run(["make", "-C", odir, "ref", f"AGBNP3_REF_DIR={ref}"], verbose)


def build(verbose=True):
    build_c(verbose)
    build_hip(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build()
