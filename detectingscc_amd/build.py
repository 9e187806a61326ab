"""Build recipe for libscc_amd.so, the `scc` executable and the CPU oracle.

Everything is built in-tree so the outputs travel with the source tree:
  detectingscc_amd/lib/libscc_amd.so   HIP kernels + host engine (gfx950)
  detectingscc_amd/bin/scc             drop-in CLI linked against it
  oracle/liboracle.so                  CPU oracle (test infrastructure)

Run `python -m detectingscc_amd.build` or call build_all().
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
INCLUDE = os.path.join(ROOT, "include")
ORACLE = os.path.join(ROOT, "oracle")

ARCH = os.environ.get("SCC_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_SOURCES = ["kernels_sweep.hip", "kernels_reach.hip", "kernels_merge.hip", "kernels_batch.hip",
               "engine.hip", "batch_engine.hip", "server.hip", "capi.hip", "measure.hip", "noalloc.hip", "gen_dev.hip"]
HOST_SOURCES = ["graph_io.cpp", "generators.cpp", "dropin.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-fvisibility=hidden",
    # no packed FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-pthread", "-fvisibility=hidden",
              f"-I{INCLUDE}", f"-I{CSRC}"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    # hipcc's host pass warns about the device-only target feature; drop that noise
    noisy = [ln for ln in r.stdout.splitlines()
             if ln.strip() and "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]
    if noisy:
        print("\n".join(noisy), file=sys.stderr)


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    else:
        _run([CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_oracle() -> str:
    _run(["make", "-s", "-C", ORACLE])
    return os.path.join(ORACLE, "liboracle.so")


def build_all(jobs: int | None = None) -> dict[str, str]:
    """Compile every HIP and host source, link the library and the CLI."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        fut_oracle = ex.submit(build_oracle)
        objs = list(ex.map(_compile, HIP_SOURCES + HOST_SOURCES))
        oracle = fut_oracle.result()
    lib = os.path.join(LIB, "libscc_amd.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-pthread", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, lib)
    exe = os.path.join(BIN, "scc")
    _run([CXX, *HOST_FLAGS, os.path.join(CSRC, "scc_cli.cpp"), "-o", exe + ".tmp",
          f"-L{LIB}", "-lscc_amd", "-Wl,-rpath,$ORIGIN/../lib", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(exe + ".tmp", exe)
    return {"lib": lib, "cli": exe, "oracle": oracle}


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    out = build_all()
    for k, v in out.items():
        print(f"{k}: {v}")
