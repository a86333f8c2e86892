"""In-tree build of the native libraries.

    libidb_hip.so    HIP kernels + C ABI (hipcc, gfx950)
    libidb_synth.so  synthetic repertoire generator (g++)
    interlink_demo   C++ drop-in InterlinkClass driver (hipcc host code)
    psdbscan         the host program: CSV + INFO config in, CSV + run log out

Run ``python -m immunodbscan_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")

HIP_SOURCES = ["idb_core.hip", "idb_pairs.hip", "idb_learned.hip", "idb_genes.hip", "idb_sort.hip", "idb_scan.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd: list[str], cwd: str | None = None) -> None:
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))


def build_hip() -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        list(ex.map(_run, jobs))
    out = os.path.join(PKG, "libidb_hip.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx",
          "-Wl,-rpath,/opt/rocm/lib", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_synth() -> str:
    out = os.path.join(PKG, "libidb_synth.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
          os.path.join(CSRC, "synth.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_interlink() -> str:
    """The drop-in C++ InterlinkClass plus a small driver that reads a
    repertoire from a binary file and writes clone ids, copies and cores."""
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(PKG, "interlink_demo")
    _run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
          os.path.join(CSRC, "Interlink.cpp"), os.path.join(CSRC, "interlink_demo.cpp"),
          "-L" + PKG, "-lidb_hip", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_psdbscan() -> str:
    """The host program over the drop-in InterlinkClass: reads CSV files and
    an INFO configuration, writes the clustered CSV and the run log."""
    out = os.path.join(PKG, "psdbscan")
    _run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
          os.path.join(CSRC, "Interlink.cpp"), os.path.join(CSRC, "psdbscan.cpp"),
          "-L" + PKG, "-lidb_hip", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_all(interlink: bool = True) -> None:
    build_synth()
    build_hip()
    if interlink and os.path.exists(os.path.join(CSRC, "Interlink.cpp")):
        build_interlink()
        build_psdbscan()


if __name__ == "__main__":
    build_all()
    print("built:", [f for f in os.listdir(PKG) if f.endswith(".so") or f in ("interlink_demo", "psdbscan")])
