"""Build libcrava_hip.so (gfx950) in-tree with hipcc.

Usage: python -m crava_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libcrava_hip.so"
HOSTLIB = PKG / "libcrava_host.so"

HIP_SOURCES = ["capi.hip", "fft.hip", "posterior.hip", "noise.hip", "simulate.hip", "synth.hip",
               "rockphysics.hip", "mlp.hip", "facies.hip", "output.hip", "krig.hip"]
HOST_SOURCES = ["host/fftgrid.cpp", "host/crava.cpp", "host/model.cpp", "host/hostapi.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay off (the compiler would otherwise form them)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{ROOT / 'include'}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    # keep warnings visible but short
    msg = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if msg:
        sys.stderr.write(msg[:4000] + "\n")


def build(jobs: int = 8, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for s in HIP_SOURCES:
        o = BUILD / (Path(s).stem + ".o")
        objs.append(o)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / s), "-o", str(o)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    # host C++ mirror of the reference classes, linked against the C ABI only
    host_srcs = [CSRC / s for s in HOST_SOURCES if (CSRC / s).exists()]
    if host_srcs:
        tmp = HOSTLIB.with_suffix(".so.tmp")
        _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-fopenmp", f"-I{ROOT / 'include'}",
              f"-I{CSRC / 'host'}", *map(str, host_srcs), "-o", str(tmp),
              f"-L{PKG}", "-lcrava_hip", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, HOSTLIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    build(a.jobs, verbose=True)


if __name__ == "__main__":
    main()
