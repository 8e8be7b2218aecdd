"""Builds the native parts of disptools_amd in-tree.

* ``libdisptools_hip.so`` -- the gfx950 kernels and the C ABI declared in
  ``include/disptools_hip.h`` (hipcc);
* ``_disptools*.so`` -- the CPython extension with the reference's method
  table (``python_c_extension/_disptools.c:532-540``), linked against it.

Usage: ``python -m disptools_amd.build`` (or ``build_all()``).
"""

import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM = os.path.dirname(os.path.dirname(HIPCC))
ARCH = os.environ.get("DISPTOOLS_OFFLOAD_ARCH", "gfx950")

LIB_NAME = "libdisptools_hip.so"
LIB_PATH = os.path.join(PKG, LIB_NAME)

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=%s" % ARCH,
    # Packed FP32 instructions are kept out of the generated code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I", INCLUDE, "-I", CSRC,
]

# Per-source contraction: the reference is ISO C99 without contraction, so the
# standalone jacobian keeps every multiply and add separately rounded (it is
# bit-exact with the reference).  The gradient pass writes each of its fused
# multiply-adds out and is compiled with contraction off, so every tile
# variant rounds every voxel the same way (the compiler could otherwise fuse
# a*b - c*d either way, differently per instantiation).
HIP_SOURCES = [("dt_kernels.hip", "-ffp-contract=off"),
               ("dt_jacobian.hip", "-ffp-contract=off"),
               ("dt_greedy_stream.hip", "-ffp-contract=fast-honor-pragmas"),
               ("dt_matching.hip", "-ffp-contract=fast-honor-pragmas"),
               ("dt_pyramid.hip", "-ffp-contract=off"),
               ("dt_noalloc.hip", "-ffp-contract=off"),
               ("dt_runtime.cpp", "-ffp-contract=off"),
               ("dt_levels.cpp", "-ffp-contract=off"),
               ("dt_engine.cpp", "-ffp-contract=off")]


# The streaming pass is compiled once per (precision, tile) variant, each
# kernel in a translation unit of its own.
# (tile x, tile y, waves per SIMD the registers are capped for)
STREAM_VARIANTS = [(ct, tx, ty, w) for ct in ("float", "double")
                   for tx, ty, w in ((32, 16, 3), (32, 20, 3), (64, 8, 3), (32, 24, 3), (64, 12, 3))]


def _stream_objects():
    out = []
    for ct, tx, ty, w in STREAM_VARIANTS:
        name = "%s_%dx%d_w%d" % ("f32" if ct == "float" else "f64", tx, ty, w)
        out.append(("dt_grad_stream.hip", "-ffp-contract=off",
                    ["-DDT_STREAM_CT=%s" % ct, "-DDT_STREAM_TX=%d" % tx, "-DDT_STREAM_TY=%d" % ty,
                     "-DDT_STREAM_MINW=%d" % w, "-DDT_STREAM_NAME=%s" % name], "dt_grad_stream_%s.o" % name))
    return out


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library(verbose=True, jobs=None):
    units = [(src, contract, [], src + ".o") for src, contract in HIP_SOURCES] + _stream_objects()
    cmds, objs = [], []
    for src, contract, defines, oname in units:
        obj = os.path.join(CSRC, oname)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([HIPCC] + HIP_FLAGS + [contract] + defines + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    jobs = jobs or max(1, min(8, os.cpu_count() or 1))
    running = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        running.append(subprocess.Popen(cmd))
        if len(running) >= jobs:
            p = running.pop(0)
            if p.wait() != 0:
                raise subprocess.CalledProcessError(p.returncode, p.args)
    for p in running:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(p.returncode, p.args)
    tmp = LIB_PATH + ".tmp"
    # roctx: named ranges per solve for rocprofv3 --marker-trace
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=%s" % ARCH] + objs +
         ["-L", os.path.join(ROCM, "lib"), "-lrocprofiler-sdk-roctx", "-Wl,-rpath," + os.path.join(ROCM, "lib"),
          "-o", tmp], verbose)
    os.replace(tmp, LIB_PATH)
    for o in objs:
        os.remove(o)
    return LIB_PATH


def extension_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(PKG, "_disptools" + suffix)


def build_extension(verbose=True):
    import numpy as np
    src = os.path.join(CSRC, "_disptools.c")
    out = extension_path()
    tmp = out + ".tmp"
    cmd = [
        "gcc", "-O2", "-std=c99", "-fPIC", "-shared", "-Wall",
        "-I", sysconfig.get_paths()["include"], "-I", np.get_include(), "-I", INCLUDE,
        src, "-o", tmp, "-L", PKG, "-ldisptools_hip", "-Wl,-rpath,$ORIGIN",
    ]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def build_all(verbose=True):
    build_library(verbose)
    build_extension(verbose)


if __name__ == "__main__":
    build_all(verbose="-q" not in sys.argv)
