"""Builds libcupic.so (HIP, gfx950) and the C++ host driver in-tree.

    python cupic_amd/build.py          # or __graft_entry__.build()

Every float expression is compiled without contraction (-ffp-contract=off)
so the kernels evaluate exactly the IEEE operations they are written with;
packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcupic.so")
DRIVER = os.path.join(HERE, "cupic_main")

SOURCES = ["k_mesh.hip", "k_particles.hip", "k_cc.hip", "k_direct.hip", "cupic_api.cpp",
           "cupic_tune.cpp", "cupic_server.cpp", "cupic_noalloc.hip", "cupic_measure.hip"]
HEADERS = ["cupic_dev.h", "cupic_launch.h"]

ARCH = os.environ.get("CUPIC_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in c or os.path.exists(c):
            return c
    return "hipcc"


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # the host pass does not know the packed-fp32 target feature; that warning is harmless
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    jobs = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [hipcc] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        jobs.append(cmd)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-o", LIB + ".tmp",
          "-Wl,--version-script=" + os.path.join(CSRC, "libcupic.map")] + objs)
    os.replace(LIB + ".tmp", LIB)
    drv = os.path.join(CSRC, "cupic_main.cpp")
    if os.path.exists(drv):
        _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), drv,
              "-o", DRIVER, "-L", HERE, "-lcupic", f"-Wl,-rpath,{HERE}",
              f"-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64"])
    return LIB


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


if __name__ == "__main__":
    build_oracle()
    print(build(verbose="-v" in sys.argv))
