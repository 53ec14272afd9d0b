"""Build of the in-tree HIP library graniitti_amd/libgra_hip.so for gfx950.

Each translation unit is compiled with hipcc; the FP64 reference-order units
(integrand, eikonal tables, host constants) are compiled without floating-point
contraction so their FP64 arithmetic rounds like the reference's.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgra_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wno-unused-result",
]
UNITS = {
    "gra_kernels_fp64.hip": ["-ffp-contract=off"],
    "gra_kernels_fp32.hip": ["-ffp-contract=fast"],
    "gra_eikonal.hip": ["-ffp-contract=off"],
    "gra_vegas.hip": ["-ffp-contract=off"],
    "gra_capi.hip": ["-ffp-contract=off"],
    "gra_flow.hip": ["-ffp-contract=fast"],
    "gra_measure.hip": ["-ffp-contract=fast"],
    "gra_config.cpp": ["-ffp-contract=off", "-x", "c++"],
    "gra_hepmc3.cpp": ["-ffp-contract=off", "-x", "c++"],
    "gra_eikcache.cpp": ["-ffp-contract=off", "-x", "c++"],
}


def _compile(unit, extra):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(BUILD, unit + ".o")
    cmd = [HIPCC] + COMMON + extra + ["-c", src, "-o", obj]
    if unit.endswith(".cpp"):
        cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
               "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed for {unit}:\n{p.stdout}\n{p.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(UNITS))) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), UNITS.items()))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stdout}\n{p.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
