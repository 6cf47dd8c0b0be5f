"""Kernel-level driver for profiling the stage kernels on one GPU.

Builds the periodic vortex workload at n^3 and times the kernels of one explicit step with
HIP events (mose_time_kernels), then runs a few full steps.  Used under rocprofv3:

    rocprofv3 --kernel-trace --stats -d <dir> -- python3 -m mose_amd.kernelbench --n 256
    rocprofv3 --pmc FETCH_SIZE -d <dir> -- python3 -m mose_amd.kernelbench --n 128 --reps 3

Options: --case vortex|tgv (tgv: --nbox viscous Taylor-Green boxes of n^3, van Leer + AUSM+),
--prec fp64|fp32, --riemann HLLC|AUSM+|HLLE|SLAU|SLAU2, --limiter vanleer|minmod|vanalbada,
--autotune picks the tile shape / k chunk / tile order by timing first,
--copy also runs a plain device-to-device copy of 5 n^3 doubles (hipMemcpy) so the HBM
byte counters (FETCH_SIZE / WRITE_SIZE) can be calibrated against a kernel of known size,
--vjp N times N residual adjoints (mose_residual_vjp) after one with Rbar = 1.
"""
import argparse
import json
import sys
import time


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="vortex", choices=["vortex", "tgv"])
    ap.add_argument("--nbox", type=int, default=8)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--prec", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--riemann", default="HLLC")
    ap.add_argument("--limiter", default="vanleer")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--copy", action="store_true", help="also time a plain copy kernel of 5 n^3 doubles")
    ap.add_argument("--autotune", action="store_true", help="run the tile autotuner first")
    ap.add_argument("--vjp", type=int, default=0,
                    help="also time N residual adjoints (mose_residual_vjp, Rbar = 1 on every interior cell)")
    args = ap.parse_args()
    from mose_amd import abi, cases
    from mose_amd.lib import GpuSolver
    prec = abi.PREC_FP64 if args.prec == "fp64" else abi.PREC_FP32_FLUX
    if args.case == "tgv":
        case = cases.taylor_green(n=args.n, nbox=args.nbox, precision=prec)
    else:
        case = cases.vortex(n=args.n, riemann=args.riemann, limiter=args.limiter, precision=prec)
    g = GpuSolver(case)
    print("[kernelbench] setup done: %d cells" % case.ncells, flush=True)
    tuned = None
    if args.autotune:
        tuned = g.autotune(3)
        print("[kernelbench] autotune: %s" % (tuned,), flush=True)
    ms = g.time_kernels(args.reps)
    t0 = time.perf_counter()
    g.advance(args.steps)
    dt = time.perf_counter() - t0
    out = {"case": args.case, "cells": case.ncells, "n": args.n, "prec": args.prec, "riemann": case.cfg.riemann,
           "limiter": case.cfg.limiter,
           "ms_dt": ms[0], "ms_ghost": ms[1], "ms_stage": ms[2], "ms_stage_last": ms[3],
           "steps": args.steps, "s_steps": dt,
           "stage_GBs_algorithmic": 120.0 * case.ncells / (ms[2] * 1e-3) / 1e9, "tiling_in_use": g.tiling(),
           "tiling": tuned[0] if tuned else None}
    if args.vjp > 0:
        import numpy as np
        Rb = [np.pad(np.ones(tuple(b.dims[::-1]) + (5,)), ((2, 2), (2, 2), (2, 2), (0, 0))) for b in case.blocks]
        g.residual_vjp(Rb)
        res_ms, vjp_ms = g.time_residual_vjp(args.vjp)[:2]
        out["vjp_residual_ms"], out["vjp_ms"], out["vjp_ratio"] = res_ms, vjp_ms, vjp_ms / res_ms
        print("[kernelbench] residual %.3f ms, vjp %.3f ms, ratio %.2f" % (res_ms, vjp_ms, vjp_ms / res_ms), flush=True)
    if args.copy:
        import ctypes as C
        hip = C.CDLL("libamdhip64.so")
        nbytes = 5 * args.n ** 3 * 8
        x, y = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(x), C.c_size_t(nbytes)) == 0
        assert hip.hipMalloc(C.byref(y), C.c_size_t(nbytes)) == 0
        assert hip.hipMemset(x, 1, C.c_size_t(nbytes)) == 0
        for _ in range(2):
            hip.hipMemcpy(y, x, C.c_size_t(nbytes), 3)
        hip.hipDeviceSynchronize()
        t0 = time.perf_counter()
        for _ in range(args.reps):
            hip.hipMemcpy(y, x, C.c_size_t(nbytes), 3)
        hip.hipDeviceSynchronize()
        out["copy_bytes_read_per_launch"] = nbytes
        out["copy_GBs"] = 2 * nbytes * args.reps / (time.perf_counter() - t0) / 1e9
        hip.hipFree(x)
        hip.hipFree(y)
    g.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
