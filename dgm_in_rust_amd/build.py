"""Builds dgm_in_rust_amd/libdgm.so (gfx950) from csrc/ with hipcc, in-tree, and beside it
libdgm_study.so, the launch-boundary study's kernels (csrc/study_sync.hip; not part of the library).

Every HIP translation unit is compiled with packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops). Objects go to dgm_in_rust_amd/_build/.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdgm.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-I" + os.path.join(HERE, "..", "include")]
HIPFLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]
SOURCES = [("kernels.hip", True), ("dgm.cpp", True), ("arena.cpp", True), ("batch.cpp", True), ("tables.cpp", False),
           ("mesh.cpp", False)]
STUDY_SOURCE = "study_sync.hip"
STUDY_LIB = os.path.join(HERE, "libdgm_study.so")


def _compile(src: str, hip: bool) -> str:
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if hip:
        cmd = [HIPCC, "-x", "hip", *HIPFLAGS, *COMMON, "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, "-x", "c++", *COMMON, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        # hipcc's host pass warns about the device-only target feature; keep other warnings visible
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")
    return obj


def _link(objs, lib: str) -> None:
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        jobs = [ex.submit(_compile, *a) for a in SOURCES]
        study = ex.submit(_compile, STUDY_SOURCE, True)
        objs = [j.result() for j in jobs]
        study_obj = study.result()
    _link(objs, LIB)
    _link([study_obj], STUDY_LIB)
    if verbose:
        print("built", LIB, "and", STUDY_LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
