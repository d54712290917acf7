"""Builds libcgs_hip.so in-tree with hipcc for gfx950.

Every HIP translation unit is compiled with packed-FP32 instructions switched
off; the setup and host translation units additionally with
-ffp-contract=off so that their FP64/FP32 sequences round exactly like the
reference's (MSVC /fp:precise) arithmetic.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcgs_hip.so")
RECO_LIB = os.path.join(HERE, "libcgs_reco.so")
RECO_MAIN = os.path.join(HERE, "cgsense_dropin")
RECO_SRC = os.path.join(CSRC, "reco")
RECO_INC = os.path.join(INCLUDE, "cgs_reco")
CXX = os.environ.get("CXX", "g++")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC, "-Wno-unused-command-line-argument",
]
SOURCES = [
    ("cgs_fft.hip", []),
    ("cgs_nufft.hip", []),
    ("cgs_cg.hip", []),
    ("cgs_setup.hip", ["-ffp-contract=off"]),
    ("cgs_voronoi.hip", ["-ffp-contract=off"]),
    ("cgs_api.hip", ["-ffp-contract=off"]),
    ("cgs_noalloc.hip", ["-ffp-contract=off"]),
    ("cgs_xlists.hip", []),
]


def _compile(src, extra):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda a: _compile(*a), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, LIB)
    build_reco()
    if verbose:
        print("built", LIB, RECO_LIB, RECO_MAIN)
    return LIB


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s\n%s" % (" ".join(cmd), r.stderr[-6000:]))


def build_reco():
    """The drop-in Reco / Fourier / IO / Trajectory classes (host C++ over the
    C-ABI) and the per-state driver program."""
    flags = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-pthread", "-ffp-contract=off", "-Wall",
             "-I" + INCLUDE, "-I" + RECO_INC]
    srcs = [os.path.join(RECO_SRC, f) for f in ("Reco.cpp", "Fourier.cpp", "IO.cpp", "Trajectory.cpp", "functions.cpp")]
    tmp = RECO_LIB + ".tmp"
    _run([CXX] + flags + ["-shared", "-o", tmp] + srcs +
         ["-L" + HERE, "-l:libcgs_hip.so", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, RECO_LIB)
    tmp = RECO_MAIN + ".tmp"
    _run([CXX] + flags + ["-o", tmp, os.path.join(RECO_SRC, "cgsense_main.cpp"),
                          "-L" + HERE, "-l:libcgs_reco.so", "-l:libcgs_hip.so", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, RECO_MAIN)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
