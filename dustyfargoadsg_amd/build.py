"""Builds libdfargo.so in-tree: hipcc for gfx950 on the HIP sources, gcc on
the host C file, no contraction anywhere (FP64 parity with the oracle)."""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libdfargo.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DFARGO_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", CSRC,
]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-Wall",
           "-I", CSRC]
HIP_SOURCES = ["step.hip", "sg.hip", "dust.hip", "api.hip", "mlp.hip"]
C_SOURCES = ["host.c", "parfile.c"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1] if cmd else ''}")
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in C_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append(["gcc"] + C_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(_run, jobs))
    tmp = OUT + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lm"])
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    build_shim(verbose)
    return OUT


SHIM = os.path.join(HERE, "shim")
SHIM_HOST = os.path.join(SHIM, "fargo_host")


def build_shim(verbose=False):
    """The drop-in shim (shim/fargo_shim.c) and the small C host that drives
    it (shim/fargo_host.c), linked against libdfargo.so with no undefined
    symbol allowed: the shim's imports are exactly what a host defines."""
    flags = ["-std=gnu11", "-O2", "-Wall", "-Wextra", "-Werror", "-ffp-contract=off",
             "-I", os.path.join(HERE, "..", "include"), "-I", SHIM]
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    pid = os.getpid()  # concurrent builds (test workers) write their own files
    for src in ("fargo_shim.c", "fargo_host.c"):
        obj = os.path.join(BUILD, src + ".o")
        tmp_obj = f"{obj}.{pid}"
        _run(["gcc"] + flags + ["-c", os.path.join(SHIM, src), "-o", tmp_obj])
        os.replace(tmp_obj, obj)
        objs.append(obj)
    tmp = f"{SHIM_HOST}.tmp{pid}"
    _run(["gcc", "-o", tmp] + objs + ["-L", HERE, "-ldfargo", "-Wl,-rpath,$ORIGIN/..", "-Wl,--no-undefined",
                                      "-lm"])
    os.replace(tmp, SHIM_HOST)
    if verbose:
        print("built", SHIM_HOST)
    return SHIM_HOST


if __name__ == "__main__":
    build(verbose=True)
