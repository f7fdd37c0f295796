"""Builds the HIP substep engine into pd_ipc_amd/libpdx.so (gfx950).

Run `python -m pd_ipc_amd.build` (or `__graft_entry__.build()`).  Every source is
compiled with hipcc; FP contraction is off so the FP64 path keeps the reference's
rounding, and packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpdx.so")
BENCH = os.path.join(HERE, "pdx_bench")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PDX_ARCH", "gfx950")

SOURCES = ["kernels.hip", "context.cpp", "factor.cpp", "amd.cpp", "nd.cpp", "scene.cpp", "mesh_io.cpp",
           "report.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-value"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
if os.environ.get("PDX_PHASE_TIMERS") == "1":
    DEVICE.append("-DPDX_PHASE_TIMERS=1")  # per-phase device timers in pdx_get_stats (slower kernels)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r.stdout


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, procs = [], []
    # every source compiles in its own process, all at once
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", "-o", obj]
        else:
            cmd = [HIPCC] + COMMON + ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-x", "c++", "-c", "-o", obj]
        cmd.append(path)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = None
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(" ".join(cmd) + "\n" + out)
            failed = failed or cmd[-1]
        elif verbose and out:
            print(out)
    if failed:
        raise RuntimeError(f"build step failed: {failed}")
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, OUT)
    # pdx_bench, the report as a program (next to the library, found through its rpath)
    _run([HIPCC, "-O2", "-std=c++17", "-x", "c++", os.path.join(CSRC, "bench_main.cpp"), "-x", "none",
          "-o", BENCH, f"-L{HERE}", "-lpdx", "-Wl,-rpath,$ORIGIN"])
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
