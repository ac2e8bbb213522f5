"""Build libcrushrs_amd.so in-tree with hipcc for gfx950.

    python -m crushrs_amd.build        (or crushrs_amd.build.build())

Objects go to crushrs_amd/_build/, the library to crushrs_amd/libcrushrs_amd.so.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libcrushrs_amd.so"
ARCH = os.environ.get("CRUSH_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
# Packed FP32 instructions stay off in device code.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _sources():
    return sorted(list(CSRC.glob("*.cpp")) + list(CSRC.glob("*.hip")))


def _sgpr_spills(remarks: str) -> dict:
    """SGPR spill counts per kernel from -Rpass-analysis=kernel-resource-usage
    remarks."""
    out, name = {}, None
    for line in remarks.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
        m = re.search(r"SGPRs Spill: (\d+)", line)
        if m and name:
            out[name] = int(m.group(1))
    return out


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    cmd = [HIPCC, *COMMON]
    if src.suffix == ".hip":
        cmd += ["-x", "hip", *DEVICE, "-Rpass-analysis=kernel-resource-usage"]
    else:
        cmd += ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
    cmd += ["-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    # the element force kernels must not spill SGPRs to VGPR lanes: a build
    # that did gave batched results that changed from run to run
    bad = {k: n for k, n in _sgpr_spills(r.stderr).items() if "k_force" in k and n > 0}
    if bad:
        raise RuntimeError(f"{src.name}: element force kernels spill SGPRs: {bad}")
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(jobs, 1)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
