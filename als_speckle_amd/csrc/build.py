"""Build libspeckle_hip.so (gfx950) in-tree with hipcc.

    python -m als_speckle_amd.csrc.build [--jobs J] [--verbose]

Each translation unit is compiled to an object under ``_build/`` next to this
file and linked into ``als_speckle_amd/libspeckle_hip.so``.  The kernel sources
are split per frame size (``SPK_ONLY_N``) so the objects build in parallel.
Always rebuilds (the caller decides when to build).
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
OUT = os.path.join(PKG, "libspeckle_hip.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SIZES = (16, 32, 64, 128, 256, 512, 1024, 2048, 4096)

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # fused multiply-adds only where the source writes fmaf: results do not
    # depend on the inlining context of the shared arithmetic
    "-ffp-contract=off",
    "-I", os.path.join(ROOT, "include"), "-I", HERE,
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
# extra -D definitions for layout experiments (space separated), e.g.
# SPK_BUILD_DEFS="-DSPK_DUAL_MIN=1024"
COMMON += os.environ.get("SPK_BUILD_DEFS", "").split()


def _units():
    units = []
    for n in SIZES:
        units.append((os.path.join(HERE, "spk_kernels.hip"),
                      os.path.join(BUILD, f"spk_kernels_{n}.o"), [f"-DSPK_ONLY_N={n}"]))
    units.append((os.path.join(HERE, "spk_kernels.hip"),
                  os.path.join(BUILD, "spk_kernels_common.o"), ["-DSPK_ONLY_N=0"]))
    units.append((os.path.join(HERE, "spk_kernels.hip"),
                  os.path.join(BUILD, "spk_kernels_group.o"), ["-DSPK_ONLY_N=-1"]))
    units.append((os.path.join(HERE, "spk_capi.hip"), os.path.join(BUILD, "spk_capi.o"), []))
    return units


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{err}")
    if err.strip() and verbose:
        print(err, file=sys.stderr)


def build(jobs: int | None = None, verbose: bool = False) -> str:
    """Compile and link; returns the path of the shared library."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    units = _units()
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC, *COMMON, *extra, "-c", src, "-o", obj], verbose)
                for src, obj, extra in units]
        for f in futs:
            f.result()
    tmp = OUT + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
          *[obj for _, obj, _ in units]], verbose)
    os.replace(tmp, OUT)
    return OUT


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    main()
