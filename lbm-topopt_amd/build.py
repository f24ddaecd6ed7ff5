"""Build recipe of liblbmx.so (HIP kernels for gfx950 + host engine), in-tree.

Objects go to lbm-topopt_amd/_build/, the library to lbm-topopt_amd/liblbmx.so.
Translation units compile in parallel; every call rebuilds (the sources are small enough).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblbmx.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    # packed FP32 instructions stay disabled on this target (see the build notes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
UNITS = ([f"kernels_{lat}_{p}.hip" for lat in ("d3q19", "d2q9", "thermal") for p in ("f64", "f32", "f16", "bf16")]
         + [f"kernels_batch_{lat}_{p}.hip" for lat in ("d2q9", "d3q19", "thermal") for p in ("f32", "f64", "f16", "bf16")]
         + ["engine.cpp", "probe.hip", "design.cpp", "design_kernels.hip", "mma.cpp", "mma_kernels.hip", "io.cpp"])


def _compile(unit):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(BUILD, os.path.splitext(unit)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    if unit.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip"] + COMMON + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stderr[-6000:]}")
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    jobs = int(os.environ.get("MAX_JOBS", "8"))
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(UNITS)))) as ex:
        objs = list(ex.map(_compile, UNITS))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


HOST = os.path.join(HERE, "host")
EXAMPLES = [os.path.join(ROOT, "tests", "cpp", name) for name in ("robust_pump_miniature.cpp", "unfused_reverse_loop.cpp")]


def build_host_examples(verbose=True):
    """Compile the C++ programs written against the host mirror (host/lbm_topopt.hpp) with the
    system g++, linked to the in-tree liblbmx.so (found at run time through a relative RUNPATH)."""
    out = []
    for src in EXAMPLES:
        exe = os.path.splitext(src)[0]
        rel = os.path.relpath(HERE, os.path.dirname(exe))
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + HOST, src,
               "-L" + HERE, "-llbmx", "-Wl,-rpath,$ORIGIN/" + rel, "-Wl,-rpath-link,/opt/rocm/lib", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host example build failed: {' '.join(cmd)}\n{r.stderr[-4000:]}")
        out.append(exe)
        if verbose:
            print(f"built {exe}")
    return out


if __name__ == "__main__":
    build()
    build_host_examples()
    sys.exit(0)
