"""Build libmose_hip.so (HIP, gfx950) and the Fortran host driver in-tree.

Usage: python -m mose_amd.build   (also called by __graft_entry__.build()).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmose_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("MOSE_FC", "/opt/rocm/llvm/bin/amdflang")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP32 division and square root of the FP32-flux kernels use the hardware reciprocal /
    # square-root instructions (~1 ulp); FP64 arithmetic stays IEEE correctly rounded
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


NLIM = 6      # limiters; mose_lim.hip is compiled once per limiter and precision
              # (MOSE_LIM_ID, MOSE_PREC_ID)
BUILD = os.path.join(HERE, "_build")


def _jobs():
    try:
        return max(1, min(int(os.environ.get("MAX_JOBS", "8")), 16))
    except ValueError:
        return 8


def build_hip(extra=()):
    """Compile the host/ABI unit and one unit per limiter and precision in parallel, then link."""
    os.makedirs(BUILD, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include")]
    units = [(os.path.join(CSRC, "mose_hip.hip"), [], os.path.join(BUILD, "mose_hip.o"))]
    units += [(os.path.join(CSRC, "mose_lim.hip"), ["-DMOSE_LIM_ID=%d" % k, "-DMOSE_PREC_ID=%d" % q],
               os.path.join(BUILD, "mose_lim%d_%d.o" % (k, q)))
              for k in range(NLIM) for q in range(2)]
    procs = []
    pending = list(units)
    failed = None
    while pending or procs:
        while pending and len(procs) < _jobs():
            src, defs, obj = pending.pop(0)
            cmd = [HIPCC] + HIP_FLAGS + list(extra) + inc + defs + ["-c", "-o", obj, src]
            print("+", " ".join(cmd), flush=True)
            procs.append((subprocess.Popen(cmd), obj))
        p, obj = procs.pop(0)
        if p.wait() != 0 and failed is None:
            failed = obj
    if failed:
        raise subprocess.CalledProcessError(1, "hipcc -> " + failed)
    tmp = LIB + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [u[2] for u in units])
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    """Fortran host driver (ISO_C_BINDING) linked against libmose_hip.so."""
    fdir = os.path.join(HERE, "fortran")
    if not os.path.exists(FC) or not os.path.exists(os.path.join(fdir, "mose_run.f90")):
        return None
    out = os.path.join(fdir, "mose_run")
    run([FC, "-O2", "-o", out, "mose_hip_iface.f90", "mose_explicit.f90", "mose_run.f90",
         "-L" + HERE, "-lmose_hip", "-Wl,-rpath,$ORIGIN/.."], cwd=fdir)
    run([FC, "-O2", "-o", os.path.join(fdir, "mose_case"), "mose_hip_iface.f90", "mose_explicit.f90",
         "mose_case.f90", "-L" + HERE, "-lmose_hip", "-Wl,-rpath,$ORIGIN/.."], cwd=fdir)
    return out


def build_oracle():
    run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def main():
    build_hip()
    build_fortran()
    build_oracle()


if __name__ == "__main__":
    sys.exit(main())
