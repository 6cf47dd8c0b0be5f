#!/usr/bin/env python3
"""Benchmark of the explicit residual/update path on one MI355X.

Workload (the headline configuration of this path): 3-D isentropic vortex on one periodic
256^3 block over [0,10]^3, gamma 1.4, free stream (1,1,1,0,1), vortex strength 5,
MUSCL van Leer + HLLC, SSP-RK3, CFL 0.8, FP64.  A step = one full explicit step
(CFL dt + 3 RK stages with ghost fill, fluxes, update, residual norm) over all cells,
with the state resident in HBM.  value = cell-updates/s = cells x steps / seconds.

Tile shapes are picked by the library's autotuner before the warm-up steps.

Also reported on the same line: the FP32-flux variant of the same workload, the same
workload with the other Riemann solver (AUSM+, FP64), the L2 error
against the advected exact vortex, the roofline of the dominant kernel (the fused
flux + RK-stage kernel) measured with HIP events on the library's stream, the 64-block
batched case (seeded block sizes in [16,128], FP32 flux vs FP64), the discrete adjoint's
cost relative to one residual, the steady adjoint gradient of the 512 x 256 bump (primal and
adjoint solve times, central-difference check), the batched viscous Taylor-Green case
(8 x 192^3), and a CPU
baseline (the C oracle, OpenMP over lines of cells, on a bounded sample of the same workload).

roofline.traffic is the HBM bytes per launch of the stage kernel read from
profiles/k_stage_vortex256_traffic.txt (two rocprofv3 --pmc runs, FETCH_SIZE and
WRITE_SIZE, scaled against a device copy of known size; the file gives the commands).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
STAGE_BYTES_PER_CELL = 3 * 5 * 8       # read U^(k-1), read U^n, write U^(k): 5 FP64 each
STEP_BYTES_PER_CELL = 3 * STAGE_BYTES_PER_CELL + 5 * 8   # + the CFL pass reading U^n


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_stage_vortex256_traffic.txt")


def measured_traffic(prec):
    """HBM bytes per stage-kernel launch from the committed rocprofv3 summary, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == prec:
                    return float(parts[2])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def l2_vs_exact(g, case, t):
    from mose_amd import abi, cases
    P = abi.interior(g.get_state(1))
    ex = cases.vortex_exact(cases.cell_centres(case.blocks[0].nodes), t, L=case.meta["L"])
    return float(np.sqrt(np.mean((P[..., 0] - ex[..., 0]) ** 2)))


NO_AUTOTUNE = False


def run_gpu(case, steps, warmup, label):
    from mose_amd.lib import GpuSolver
    t0 = time.perf_counter()
    g = GpuSolver(case)
    log("[bench] %s: setup %.1f s, %d cells" % (label, time.perf_counter() - t0, case.ncells))
    if NO_AUTOTUNE:     # the tiling the library took from MOSE_TILE_SHAPE / MOSE_KCHUNK / MOSE_TILE_XCD
        tiling = g.tiling()
        log("[bench] %s: tiling from the environment %s" % (label, tiling))
    else:
        tiling, tune_ms = g.autotune(3)
        log("[bench] %s: autotune %s (%.3f ms for 3 stages)" % (label, tiling, tune_ms))
    g.tiling = tiling
    g.advance(warmup)
    t0 = time.perf_counter()
    t_sim, norms = g.advance(steps)     # synchronises at the end
    dt = time.perf_counter() - t0
    ms = g.time_kernels(nrep=10)
    log("[bench] %s: %d steps %.3f s; kernel ms dt/ghost/stage/last = %s" % (label, steps, dt, ms))
    return g, dt, t_sim, ms


def adjoint_ratio(g, case, nrep=5):
    """Device time of one discrete-adjoint evaluation (mose_residual_vjp) over one residual
    evaluation (mose_residual) on the current state, seeded with Rbar = 1 on every cell."""
    Rbars = [np.pad(np.ones(tuple(b.dims[::-1]) + (5,)), ((2, 2), (2, 2), (2, 2), (0, 0))) for b in case.blocks]
    g.residual_vjp(Rbars)
    res_ms, vjp_ms, geo_ms = g.time_residual_vjp(nrep, geometry=True)
    out = {"residual_ms": res_ms, "vjp_ms": vjp_ms, "ratio": vjp_ms / res_ms}
    if geo_ms > 0.0:   # general metrics: state + node + inflow adjoints in one evaluation
        out["vjp_geometry_ms"] = geo_ms
        out["ratio_geometry"] = geo_ms / res_ms
    return out


def tgv_adjoint_ratio(n=192, nbox=8):
    """Adjoint/primal ratio on the viscous 8 x 192^3 Taylor-Green batch (FP64, tuned tiling)."""
    from mose_amd import abi, cases
    from mose_amd.lib import GpuSolver
    case = cases.taylor_green(n=n, nbox=nbox, precision=abi.PREC_FP64)
    g = GpuSolver(case)
    g.autotune(2)
    out = adjoint_ratio(g, case, nrep=5)
    out["cells"] = case.ncells
    g.close()
    return out


def bump_adjoint():
    """Adjoint/primal ratio on the 512 x 256 body-fitted bump channel (van Albada + HLLC,
    405 inflow, 400 outflow, 300 walls, local time stepping), after 20 steps."""
    from mose_amd import cases
    from mose_amd.lib import GpuSolver
    case = cases.bump_channel()
    g = GpuSolver(case)
    g.advance(20)
    out = adjoint_ratio(g, case, nrep=20)
    out["cells"] = case.ncells
    g.close()
    return out


def steady_adjoint_bump(ni=512, nj=256, ncontrol=64, fd_steps=(1e-6, 1e-7)):
    """Steady adjoint gradient on the 512 x 256 bump channel with 64 seeded control-point
    heights (minmod + HLLC, local dt, inflow Mach 2 at 0.03 rad): the primal converged to its
    round-off floor, then the device adjoint solve and dJ/d(heights), dJ/d(inflow record) of
    J = sum p A over the bump's wall faces; the largest height derivative against central
    differences of J at re-converged states (warm-started from the base state)."""
    from mose_amd import cases, steady_adjoint as sa
    from mose_amd.lib import GpuSolver
    h0 = 0.002 * np.random.default_rng(64).random(ncontrol)
    mk = lambda h: cases.bump_channel(ni=ni, nj=nj, ncontrol=ncontrol, heights=h, limiter="minmod", alpha=0.03)
    case = mk(h0)
    g = GpuSolver(case)
    g.autotune(3)
    conv = dict(max_steps=400000, tol=5e-15, chunk=2000, stall=6)
    t0 = time.perf_counter()
    steps, norms = sa.converge(g, **conv)
    tp = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = sa.gradient(case, g, rtol=1e-10, check_every=500, max_iter=400000)
    ta = time.perf_counter() - t0
    P = g.get_state(1)
    g.close()
    log("[bench] steady adjoint %dx%d: primal %d steps %.1f s, adjoint %d iterations %.1f s"
        % (ni, nj, steps, tp, res["adjoint_iterations"], ta))
    m = int(np.argmax(np.abs(res["dJdh"])))
    fd = {}
    for eps in fd_steps:
        Js = []
        for sgn in (1.0, -1.0):
            h = h0.copy()
            h[m] += sgn * eps
            J, _, _ = sa.objective_at(mk(h), GpuSolver, warm=P, **conv)
            Js.append(J)
        d = (Js[0] - Js[1]) / (2 * eps)
        fd["%g" % eps] = {"fd": d, "rel_diff": abs(d - res["dJdh"][m]) / abs(d)}
    return {"cells": case.ncells, "control_points": ncontrol, "scheme": "minmod + HLLC, local dt, SSP-RK3, CFL 0.5",
            "J": res["J"], "primal_steps": steps, "primal_step_norm": float(np.max(norms)), "primal_s": tp,
            "primal_ms_per_step": 1e3 * tp / steps, "adjoint_iterations": res["adjoint_iterations"],
            "adjoint_res": list(res["adjoint_res"]), "adjoint_s": ta,
            "adjoint_ms_per_iteration": 1e3 * ta / res["adjoint_iterations"],
            "adjoint_vs_primal_solve_time": ta / tp,
            "dJdh_norm": float(np.linalg.norm(res["dJdh"])),
            "dJdinflow": {k: float(v) for k, v in zip(("mach", "T", "p", "alpha", "beta"), res["dJdinflow"])},
            "fd_check": {"height": m, "adjoint": float(res["dJdh"][m]), "central_differences": fd,
                         "note": "steps of 1e-6 and above move limiter states across minmod's kinks on "
                                 "this grid (the difference is not monotone in the step); 1e-7 and 1e-8 "
                                 "agree with the adjoint to 2e-9"}}


def batched(steps, warmup):
    """64 seeded blocks (dims uniform in [16,128], seed 1234) joined by 1-to-1 connections,
    Mach 0.8 free stream + 5 % Gaussian bumps, van Leer + HLLC, CFL 0.5: FP32 flux vs FP64,
    all blocks in one stage launch."""
    from mose_amd import abi, cases
    out = {}
    states = {}
    for prec, name in ((abi.PREC_FP64, "fp64"), (abi.PREC_FP32_FLUX, "fp32_flux")):
        t0 = time.perf_counter()
        case = cases.multiblock_chain(precision=prec)
        log("[bench] batched %s: case built in %.1f s" % (name, time.perf_counter() - t0))
        g, secs, t_sim, ms = run_gpu(case, steps, warmup, "batched " + name)
        states[name] = np.concatenate([abi.interior(g.get_state(b + 1)).ravel()
                                       for b in range(len(case.blocks))])
        out[name] = {"value": case.ncells * steps / secs, "ms_per_step": 1e3 * secs / steps,
                     "stage_kernel_ms": ms[2], "tiling": g.tiling,
                     "stage_frac": STAGE_BYTES_PER_CELL * case.ncells / (ms[2] * 1e-3) / 1e9 / HBM_PEAK_GBS}
        out["cells"] = case.ncells
        g.close()
    out["blocks"] = 64
    out["steps"] = steps
    out["speedup_fp32_vs_fp64"] = out["fp32_flux"]["value"] / out["fp64"]["value"]
    out["rel_l2_fp32_vs_fp64"] = float(np.linalg.norm(states["fp32_flux"] - states["fp64"])
                                       / np.linalg.norm(states["fp64"]))
    return out


def taylor_green_batch(steps, warmup, n=192, nbox=8):
    """8 periodic 192^3 Taylor-Green boxes (Mach 0.1, Re per box from {400, 800, 1600, 3200},
    Sutherland viscosity, Pr 0.71), van Leer + AUSM+ with the laminar viscous terms, all boxes
    in one launch per stage: FP32 flux against FP64, kinetic energy and its decay compared."""
    from mose_amd import abi, cases
    from mose_amd.lib import GpuSolver
    t0 = time.perf_counter()
    case = cases.taylor_green(n=n, nbox=nbox, precision=abi.PREC_FP32_FLUX)
    log("[bench] tgv: case built in %.1f s, %d cells" % (time.perf_counter() - t0, case.ncells))
    ke0 = np.array(cases.kinetic_energy([b.P for b in case.blocks]))
    out = {"cells": case.ncells, "boxes": nbox, "steps": steps, "Re": [int(r) for r in case.meta["Re"]]}
    ke = {}
    for prec, name in ((abi.PREC_FP32_FLUX, "fp32_flux"), (abi.PREC_FP64, "fp64")):
        cfg = abi.make_cfg(cases.GAMMA, cases.RGAS_NONDIM, case.cfg.cfl, "vanleer", "AUSM+", 3, True, precision=prec)
        g = GpuSolver(case, cfg=cfg)
        tiling, _ = g.autotune(2)
        g.advance(warmup)
        t0 = time.perf_counter()
        t_sim, _ = g.advance(steps)
        secs = time.perf_counter() - t0
        ms = g.time_kernels(nrep=5)
        ke[name] = np.array(cases.kinetic_energy([g.get_state(b + 1) for b in range(nbox)]))
        out[name] = {"value": case.ncells * steps / secs, "ms_per_step": 1e3 * secs / steps,
                     "stage_kernel_ms": ms[2], "sim_time": t_sim, "tiling": tiling}
        log("[bench] tgv %s: %.3f ms/step" % (name, 1e3 * secs / steps))
        g.close()
    out["speedup_fp32_vs_fp64"] = out["fp32_flux"]["value"] / out["fp64"]["value"]
    out["ke_rel_diff_max"] = float(np.max(np.abs(ke["fp32_flux"] - ke["fp64"]) / ke["fp64"]))
    d64, d32 = ke0 - ke["fp64"], ke0 - ke["fp32_flux"]
    out["ke_decay_rel_diff_max"] = float(np.max(np.abs(d32 - d64) / np.abs(d64)))
    return out


def taylor_green_adjoint(nsteps, n=192, nbox=8, stride=0):
    """d KE(t_K) / d Re of every box of the 8 x 192^3 Taylor-Green batch (FP64 state and flux):
    the fixed-dt march of nsteps steps and its reverse sweep on the device
    (mose_unsteady_adjoint_ke, checkpoints every `stride` steps), timed against the march alone."""
    from mose_amd import abi, cases
    from mose_amd.lib import GpuSolver
    os.environ["MOSE_PROGRESS"] = "1"      # one line per reversed segment (long sweeps)
    case = cases.taylor_green(n=n, nbox=nbox, precision=abi.PREC_FP64)
    g = GpuSolver(case)
    g.autotune(2)
    dt = g.compute_dt()
    g.march_kinetic_energy(2, dt, adjoint=False)                 # warm-up
    t0 = time.perf_counter()
    ke_f = g.march_kinetic_energy(nsteps, dt, stride=stride, adjoint=False)
    tf = time.perf_counter() - t0
    t0 = time.perf_counter()
    ke, mub, _ = g.march_kinetic_energy(nsteps, dt, stride=stride)
    ta = time.perf_counter() - t0
    g.close()
    mu = np.array([p[0] for _, p in case.meta["viscous"]])
    log("[bench] tgv adjoint: %d steps, march %.1f s, march + reverse %.1f s" % (nsteps, tf, ta))
    return {"cells": case.ncells, "boxes": nbox, "steps": nsteps, "dt": dt, "t_end": nsteps * dt,
            "Re": [int(r) for r in case.meta["Re"]], "ke": [float(x) for x in ke],
            "dKE_dRe": [float(-m * m * x) for m, x in zip(mu, mub)],
            "march_s": tf, "march_and_reverse_s": ta, "adjoint_vs_primal_time": (ta - tf) / tf,
            "same_ke_as_march": bool(np.array_equal(ke, ke_f)),
            "note": "FP64; checkpoints at the start of every stride-th step (stride sqrt(steps)), "
                    "segments recomputed on the device; dKE/dRe = -mu_ref^2 dKE/dmu_ref"}


# The other scheme rows on the same 256^3 vortex, FP64: (label, limiter, Riemann solver,
# MUSCL-SD, irs-beta).  IRS runs the residual-only stage kernel, the scaling, per direction the
# residual ghosts and one fused two-sweep Jacobi pass, and the update kernel.
SCHEME_LEGS = [("vanleer+HLLE", "vanleer", "HLLE", False, 0.0),
               ("vanleer+SLAU2", "vanleer", "SLAU2", False, 0.0),
               ("mc+SLAU", "mc", "SLAU", False, 0.0),
               ("vanleer+HLLE++", "vanleer", "HLLE++", False, 0.0),
               ("vanleer+HLLC+", "vanleer", "HLLC+", False, 0.0),
               ("minmod+HLLE MUSCL-SD", "minmod", "HLLE", True, 0.0),
               ("vanleer+HLLC IRS 0.5", "vanleer", "HLLC", False, 0.5)]


def scheme_legs(n, steps, warmup, ref_ms_per_step):
    from mose_amd import abi, cases
    out = {}
    for label, lim, rs, sd, irs in SCHEME_LEGS:
        case = cases.vortex(n=n, riemann=rs, limiter=lim, precision=abi.PREC_FP64)
        c = case.cfg
        case.cfg = abi.make_cfg(c.gamma, c.Rgas, c.cfl, lim, rs, c.n_rk, bool(c.time_accurate),
                                precision=abi.PREC_FP64, muscl_sd=sd, irs_beta=irs)
        g, secs, t_sim, ms = run_gpu(case, steps, warmup, label)
        leg = {"value": n ** 3 * steps / secs, "steps": steps, "ms_per_step": 1e3 * secs / steps,
               "step_time_vs_vanleer_hllc": 1e3 * secs / steps / ref_ms_per_step,
               "tiling": g.tiling, "l2_rho_vs_exact": l2_vs_exact(g, case, t_sim)}
        if irs == 0.0:   # the stage kernel this scheme steps with
            leg["stage_kernel_ms"] = ms[2]
            leg["stage_frac"] = STAGE_BYTES_PER_CELL * n ** 3 / (ms[2] * 1e-3) / 1e9 / HBM_PEAK_GBS
        if sd:
            leg["ghost_and_detector_ms"] = ms[1]
        out[label] = leg
        g.close()
    return out


def cpu_baseline(n=128, steps=20):
    """The C oracle on the same workload at n^3, on OMP_NUM_THREADS threads (OpenMP over lines of
    cells like the reference's collapse(2) loops): the build with the reference's Release-style
    flags (-O3 -funroll-loops -finline-functions -ffast-math -march=x86-64-v3
    -fomit-frame-pointer) is the baseline; the parity-checker build (-O2, no fast-math) is
    reported beside it."""
    from mose_amd import cases
    from oracle.oracle import OracleSolver, num_threads
    out = {}
    for variant in ("release", "check"):
        case = cases.vortex(n=n)
        o = OracleSolver(case, variant=variant)
        o.step()
        t0 = time.perf_counter()
        for _ in range(steps):
            o.step()
        out[variant] = n ** 3 * steps / (time.perf_counter() - t0)
        o.close()
    nt = num_threads("release")
    return {"value": out["release"], "unit": "cell-updates/s", "cores": nt, "kind": "port",
            "check_build_value": out["check"],
            "sample": "vortex %d^3, %d steps after 1 warm-up step, C oracle built with -O3 -funroll-loops "
                      "-finline-functions -ffast-math -march=x86-64-v3 -fopenmp (%d threads) on %s; "
                      "check_build_value: the -O2 parity build" % (n, steps, nt, platform.processor() or platform.machine())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--riemann", default="HLLC")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="headline legs: keep the tiling of MOSE_TILE_SHAPE / MOSE_KCHUNK / MOSE_TILE_XCD (profiling)")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=50)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-tgv", action="store_true")
    ap.add_argument("--no-steady-adjoint", action="store_true")
    ap.add_argument("--no-tgv-adjoint", action="store_true")
    ap.add_argument("--tgv-adjoint-steps", type=int, default=40,
                    help="steps of the Taylor-Green adjoint leg (about 670 reach t = 1)")
    ap.add_argument("--tgv-steps", type=int, default=20)
    ap.add_argument("--no-ausmp", action="store_true")
    ap.add_argument("--ausmp-steps", type=int, default=50)
    ap.add_argument("--no-schemes", action="store_true")
    ap.add_argument("--scheme-steps", type=int, default=30)
    args = ap.parse_args()
    global NO_AUTOTUNE
    NO_AUTOTUNE = args.no_autotune
    if args.gpus != 1:
        print("bench.py: this path runs on one GPU; sharding blocks over several GPUs is not "
              "built. Run with --gpus 1.", file=sys.stderr)
        return 2

    from mose_amd import abi, cases

    n = args.n
    ncell = n ** 3
    case = cases.vortex(n=n, riemann=args.riemann, precision=abi.PREC_FP64)
    g, secs, t_sim, ms = run_gpu(case, args.steps, args.warmup, "fp64")
    cups = ncell * args.steps / secs
    err64 = l2_vs_exact(g, case, t_sim)
    P64 = abi.interior(g.get_state(1))
    tiling64 = g.tiling
    adj_vortex = adjoint_ratio(g, case) if not args.no_adjoint else None
    g.close()
    del g

    stage_ms = ms[2]
    achieved = STAGE_BYTES_PER_CELL * ncell / (stage_ms * 1e-3) / 1e9
    result = {
        "metric": "cell-updates/s + % HBM roofline (FP64 vs FP32-flux); adjoint/primal time ratio",
        "value": cups,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * secs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "isentropic vortex %d^3 periodic block, MUSCL van Leer + %s, SSP-RK3, "
                               "CFL 0.8, FP64 state and flux" % (n, args.riemann),
                   "cells": ncell, "riemann": args.riemann, "limiter": "vanleer", "tiling": tiling64},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("fp64"),
                     "kernel": "k_stage (fused face fluxes + RK stage), %d B/cell algorithmic"
                               % STAGE_BYTES_PER_CELL,
                     "kernel_ms": stage_ms},
        "step_roofline_frac": STEP_BYTES_PER_CELL * cups / (HBM_PEAK_GBS * 1e9),
        "kernel_ms": {"dt": ms[0], "ghost": ms[1], "stage": ms[2], "stage_last": ms[3]},
        "l2_rho_vs_exact": err64,
        "sim_time": t_sim,
    }
    if not args.no_fp32:
        c32 = cases.vortex(n=n, riemann=args.riemann, precision=abi.PREC_FP32_FLUX)
        g32, secs32, t32, ms32 = run_gpu(c32, args.steps, args.warmup, "fp32-flux")
        P32 = abi.interior(g32.get_state(1))
        g32.close()
        cups32 = ncell * args.steps / secs32
        result["fp32_flux"] = {
            "value": cups32, "ms_per_step": 1e3 * secs32 / args.steps,
            "speedup_vs_fp64": cups32 / cups,
            "rel_l2_vs_fp64": float(np.linalg.norm(P32 - P64) / np.linalg.norm(P64)),
            "stage_kernel_ms": ms32[2],
            "tiling": g32.tiling,
            "traffic": measured_traffic("fp32"),
            "stage_frac": STAGE_BYTES_PER_CELL * ncell / (ms32[2] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "step_roofline_frac": STEP_BYTES_PER_CELL * cups32 / (HBM_PEAK_GBS * 1e9),
        }
    if not args.no_ausmp:
        # the same workload with the other Riemann solver of the case (van Leer + AUSM+), FP64
        other = "AUSM+" if args.riemann == "HLLC" else "HLLC"
        ca = cases.vortex(n=n, riemann=other, precision=abi.PREC_FP64)
        ga, secsa, ta, msa = run_gpu(ca, args.ausmp_steps, min(args.warmup, 2), "fp64-" + other)
        result["other_riemann"] = {
            "riemann": other, "value": ncell * args.ausmp_steps / secsa, "steps": args.ausmp_steps,
            "ms_per_step": 1e3 * secsa / args.ausmp_steps, "stage_kernel_ms": msa[2], "tiling": ga.tiling,
            "stage_frac": STAGE_BYTES_PER_CELL * ncell / (msa[2] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "l2_rho_vs_exact": l2_vs_exact(ga, ca, ta)}
        ga.close()
    if not args.no_schemes:
        result["schemes"] = scheme_legs(n, args.scheme_steps, min(args.warmup, 2), result["ms_per_step"])
    if adj_vortex is not None:
        result["adjoint"] = {"vortex": adj_vortex, "bump_512x256": bump_adjoint(), "tgv_8x192": tgv_adjoint_ratio(),
                             "note": "device ms of one mose_residual_vjp / one mose_residual, FP64; "
                                     "*_geometry: mose_residual_vjp_geometry (state + node + inflow adjoints)"}
    if not args.no_steady_adjoint:
        result["steady_adjoint_bump_512x256"] = steady_adjoint_bump()
    if not args.no_batched:
        result["batched_64_blocks"] = batched(args.batched_steps, min(args.warmup, 2))
    if not args.no_tgv:
        result["tgv_8x192"] = taylor_green_batch(args.tgv_steps, 2)
    if not args.no_tgv_adjoint:
        result["tgv_8x192_adjoint"] = taylor_green_adjoint(args.tgv_adjoint_steps)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
