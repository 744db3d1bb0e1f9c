"""Build the native library ``hyquas_amd/libhyquas_amd.so``, the CLI
``hyquas_amd/hyquas_main`` and the kernel compile helper ``hyquas_amd/hq_rtc``
in-tree.

Device sources (``csrc/device/*.hip``) are compiled by hipcc for gfx950 only;
host sources (``csrc/host/*.cpp``) by g++ with OpenMP; everything is linked
by hipcc.  Objects go to ``hyquas_amd/_build``.  Always recompiles.

    python -m hyquas_amd.build [-j N] [-v]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhyquas_amd.so")
CLI = os.path.join(HERE, "hyquas_main")
RTC = os.path.join(HERE, "hq_rtc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX_HOST", "g++")

INCLUDES = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(CSRC, "host"),
            "-I" + os.path.join(CSRC, "device")]
# packed-FP32 instructions are disabled for every HIP file (-packed-fp32-ops)
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"] + INCLUDES
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-fopenmp"] + INCLUDES


def _sources():
    dev = sorted(os.path.join(CSRC, "device", f) for f in os.listdir(os.path.join(CSRC, "device"))
                 if f.endswith(".hip"))
    host = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host"))
                  if f.endswith(".cpp") and f != "main.cpp")
    return dev, host


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops' is not a recognized" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out, flush=True)
    return out


def _obj(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    return os.path.join(BUILD, rel + ".o")


def build(jobs: int | None = None, verbose: bool = False) -> str:
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found; the HIP toolchain is required to build hyquas_amd")
    os.makedirs(BUILD, exist_ok=True)
    dev, host = _sources()
    cmds = []
    for s in dev:
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", s, "-o", _obj(s)])
    for s in host:
        cmds.append([CXX] + CXX_FLAGS + ["-c", s, "-o", _obj(s)])
    main_src = os.path.join(CSRC, "host", "main.cpp")
    cmds.append([CXX] + CXX_FLAGS + ["-c", main_src, "-o", _obj(main_src)])
    # the compile helper: plain C++ over hiprtc, no GPU code
    cmds.append([CXX, "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
                 "-I" + os.path.join(CSRC, "device"), os.path.join(CSRC, "rtc", "hq_rtc.cpp"),
                 "-L" + os.path.join(ROCM, "lib"), "-lhiprtc", "-Wl,-rpath," + os.path.join(ROCM, "lib"),
                 "-o", RTC + ".tmp"])
    jobs = jobs or min(8, os.cpu_count() or 1)
    # longest first: the OShareMem kernels dominate
    cmds.sort(key=lambda c: 0 if "sharemem" in c[-3] else 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    os.replace(RTC + ".tmp", RTC)
    objs = [_obj(s) for s in dev + host]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-lgomp", "-lhiprtc", "-Wl,-soname,libhyquas_amd.so"], verbose)
    os.replace(tmp, LIB)
    _run([HIPCC, "--offload-arch=" + ARCH, "-o", CLI, _obj(main_src), "-L" + HERE, "-lhyquas_amd",
          "-Wl,-rpath,$ORIGIN", "-lgomp"], verbose)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    print(build(a.j, a.v))


if __name__ == "__main__":
    sys.exit(main())
