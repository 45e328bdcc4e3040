"""Build recipes for the native libraries, in-tree under epigpu_amd/lib/.

  libepiscan.so  hipcc, gfx950 only: HIP kernels + the episcan.h C ABI
  libepihost.so  gcc: host driver, codec, container, PLINK ingest, generator, F helpers
  episcan        gcc: analysis-mode CLI linked to both
The oracle (oracle/Makefile) is test infrastructure and is built separately.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "epigpu_amd")
LIB = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
HOST_SRC = ["codec.c", "synth.c", "fdist.c", "container.c", "plink.c", "perform_scan.c", "compat.c"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    if verbose and out:
        print(out)


def build_episcan(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libepiscan.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I" + INC, os.path.join(PKG, "csrc", "episcan.hip"), "-o", tmp], verbose)
    os.replace(tmp, out)
    return out


def build_epihost(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libepihost.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(PKG, "host", s) for s in HOST_SRC]
    _run([CC, "-O2", "-std=gnu11", "-Wall", "-fPIC", "-shared", "-fopenmp", "-I" + INC, *srcs,
          "-L" + LIB, "-lepiscan", "-Wl,-rpath,$ORIGIN", "-lm", "-o", tmp], verbose)
    os.replace(tmp, out)
    exe = os.path.join(LIB, "episcan")
    _run([CC, "-O2", "-std=gnu11", "-Wall", "-I" + INC, os.path.join(PKG, "host", "episcan_main.c"),
          "-L" + LIB, "-lepihost", "-lepiscan", "-Wl,-rpath,$ORIGIN", "-lm", "-o", exe + ".tmp"], verbose)
    os.replace(exe + ".tmp", exe)
    return out


def build_oracle(verbose=False):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], verbose)
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "dropin"], verbose)


def build_all(verbose=False):
    build_episcan(verbose)
    build_epihost(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
