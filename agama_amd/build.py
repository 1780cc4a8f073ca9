"""Build libagama_amd.so in-tree: host construction code with g++, HIP kernels and the C ABI
with hipcc for gfx950.  Usage: python agama_amd/build.py  (runs without importing the package)."""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.environ.get("AGX_BUILD_OUT", os.path.join(HERE, "libagama_amd.so"))
BUILD = os.environ.get("AGX_BUILD_DIR", os.path.join(HERE, "_build"))
ARCH = os.environ.get("AGX_OFFLOAD_ARCH", "gfx950")

HOST_FLAGS = ["-std=c++17", "-O2", "-march=x86-64-v3", "-ffp-contract=off", "-fPIC", "-Wall"]
HIP_FLAGS = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"] + os.environ.get("AGX_EXTRA_FLAGS", "").split()


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


HOST_SOURCES = ("host_multipole.cpp", "host_lyapunov.cpp")
HIP_SOURCES = ("agx_kernels.hip", "agx_orbit_rkn.hip", "agx_hessian.hip", "agx_orbit_var.hip", "agx_orbit_interp.hip",
               "agx_construct.hip", "agx_schedule.hip", "agx_surrogate.hip", "agx_orbit_surrogate.hip",
               "agx_capi.hip")


def build(verbose_resources=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmds, objs = [], []
    for src in HOST_SOURCES:
        o = os.path.join(BUILD, src.replace(".cpp", ".o"))
        cmds.append(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in HIP_SOURCES:
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        extra = ["-Rpass-analysis=kernel-resource-usage"] if verbose_resources and src != "agx_capi.hip" else []
        cmds.append([hipcc] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    # the translation units are independent: compile them side by side (at most 8 at a time)
    jobs = max(1, min(int(os.environ.get("AGX_BUILD_JOBS", "0")) or (os.cpu_count() or 1), 8))
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(_run, c) for c in cmds]:
            f.result()
    tmp = OUT + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose_resources="-v" in sys.argv)
