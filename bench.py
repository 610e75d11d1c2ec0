#!/usr/bin/env python3
"""Benchmark of the FLIP/APIC substep on one MI355X.

Workload: the Ultra sealed hourglass (896x504 cells, 1.59 mm, ~0.9 M particles), frames of
1/480 s split into substeps by the reference's adaptive bound, pressure solved by MGPCG to a
relative residual of 1e-6.  One bench step = one frame (air update + its substeps).  The
mixed-precision run (FP16 multigrid storage, FP32 PCG, FP64 dots) is the reported value; the
FP32-multigrid run replays the same substep schedule for the comparison.

Prints ONE JSON line (see the keys below).  --gpus N > 1 is not supported (one GPU per box).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tanque-de-agua_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DOMINANT_KERNEL = "mg_up_l0"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(tq, work, precision, warmup, steps, schedule=None, time_kernel=None, tol=1e-6, maxit=200):
    from tanque_hip.scenes import passo_frame
    cena = work.cena()
    sim = tq.SimulacaoHIP(cena, precision=precision, tol=tol, iteracoes=maxit)
    sim.adicionar(work.particles, False)
    n0 = sim.n
    sched = []
    h, H, dt = work.h, cena.altura, work.dt_frame

    def frame(k):
        sub = schedule[k] if schedule is not None else passo_frame(h, H, sim.vMax, dt)
        sched.append(sub)
        sim.avancar(dt, sub)
        return sub

    for k in range(warmup):
        frame(k)
    sim.sync()
    st0 = sim.solver_stats()
    if time_kernel:
        sim.timing_enable(True, only=time_kernel)
    t0 = time.perf_counter()
    subs = 0
    for k in range(warmup, warmup + steps):
        subs += frame(k)
    sim.sync()
    t1 = time.perf_counter()
    timing = sim.timing_read() if time_kernel else {}
    if time_kernel:
        sim.timing_enable(False)
    st1 = sim.solver_stats(true_residual=True)
    n1 = sim.n
    iters = st1["total_iterations"] - st0["total_iterations"]
    solves = st1["solves"] - st0["solves"]
    res = {
        "seconds": t1 - t0, "substeps": subs, "substeps_per_s": subs / (t1 - t0),
        "particles_start": n0, "particles_end": n1,
        "particle_steps_per_s": subs * 0.5 * (n0 + n1) / (t1 - t0),
        "pcg_iters_per_solve": iters / max(solves, 1),
        "last_true_rel_residual": st1["true_resid"] / st1["norm_b"] if st1["norm_b"] > 0 else 0.0,
        "last_rel_residual": st1["norm"] / st1["norm_b"] if st1["norm_b"] > 0 else 0.0,
        "active_bubbles": st1["active_bubbles"], "levels": st1["levels"],
        "schedule": sched, "timing": timing,
    }
    sim.close()
    return res


def cpu_baseline(work, budget_s=15.0, tol=1e-6, maxit=200):
    """The oracle (single-threaded C++ restatement of the reference CPU engine) on a bounded
    sample of the same workload: the first frames of the same scene, same particles."""
    from oracle import oracle as orc
    from tanque_hip.scenes import passo_frame
    orc.build()
    cena = orc.Cena(work.nx, work.ny, work.h)
    for s in work.strokes:
        cena.tracar(*s)
    sim = orc.Simulacao(cena, mode=0, tol=tol, iteracoes=maxit)
    sim.adicionar(work.particles, False)
    t0 = time.perf_counter()
    subs = frames = 0
    while time.perf_counter() - t0 < budget_s and frames < work.frames:
        sub = int(math.ceil(work.dt_frame / orc.passo_maximo(work.h, cena.altura, sim.vmax, work.dt_frame)))
        sim.avancar(work.dt_frame, sub)
        subs += sub
        frames += 1
    el = time.perf_counter() - t0
    return {"value": subs / el, "unit": "substeps/s", "cores": 1, "kind": "port",
            "sample": "first %d frames (%d substeps) of the same workload, oracle JS-exact mode, tol %g, "
                      "%.1f s" % (frames, subs, tol, el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120, help="timed frames")
    ap.add_argument("--warmup", type=int, default=10, help="untimed frames")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--nx", type=int, default=896)
    ap.add_argument("--ny", type=int, default=504)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import tanque_hip as tq
    from tanque_hip import scenes
    tq.load()
    work = scenes.ultra_hourglass() if (args.nx, args.ny) == (896, 504) else scenes.scaled_hourglass(args.nx, args.ny)
    log("workload %s: %dx%d cells, %d particles seeded" % (work.name, work.nx, work.ny, len(work.particles) // 12))

    mixed = run_gpu(tq, work, tq.PREC_MIXED, args.warmup, args.steps, time_kernel=DOMINANT_KERNEL)
    log("mixed: %.1f substeps/s, %.2f PCG it/solve, last rel resid %.2e (true %.2e)" % (
        mixed["substeps_per_s"], mixed["pcg_iters_per_solve"], mixed["last_rel_residual"],
        mixed["last_true_rel_residual"]))
    fp32 = None
    if not args.no_fp32:
        fp32 = run_gpu(tq, work, tq.PREC_FP32, args.warmup, args.steps, schedule=mixed["schedule"])
        log("fp32:  %.1f substeps/s, %.2f PCG it/solve, last rel resid %.2e" % (
            fp32["substeps_per_s"], fp32["pcg_iters_per_solve"], fp32["last_rel_residual"]))

    kt = mixed["timing"].get(DOMINANT_KERNEL)
    roofline = None
    if kt and kt["launches"]:
        avg_s = kt["ms"] / kt["launches"] * 1e-3
        achieved = kt["bytes_per_launch"] / avg_s / 1e9
        roofline = {"bound": "hbm", "kernel": DOMINANT_KERNEL, "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": None, "avg_launch_us": round(avg_s * 1e6, 3),
                    "bytes_per_launch": kt["bytes_per_launch"], "launches": kt["launches"]}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(work, args.cpu_budget)
        log("cpu baseline: %.3f substeps/s (%s)" % (cpu["value"], cpu["sample"]))

    ms_per_step = mixed["seconds"] / args.steps * 1e3
    line = {
        "metric": "substeps/s @~450k cells, mixed vs FP32 MGPCG at 1e-6 resid; % HBM peak",
        "value": round(mixed["substeps_per_s"], 2),
        "unit": "substeps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f16 MG storage, f64 dots",
        "data": "synthetic",
        "config": {"workload": "%s: %s, frames of %.6f s, PCG to rel. residual 1e-6" % (
            work.name, work.description, work.dt_frame),
            "cells": work.nx * work.ny, "particles": mixed["particles_start"],
            "substeps_timed": mixed["substeps"], "pcg_iters_per_solve": round(mixed["pcg_iters_per_solve"], 3),
            "last_rel_residual": mixed["last_rel_residual"], "last_true_rel_residual": mixed["last_true_rel_residual"],
            "active_bubbles_end": mixed["active_bubbles"], "mg_levels": mixed["levels"]},
        "particle_steps_per_s": round(mixed["particle_steps_per_s"], 1),
        "fp32_substeps_per_s": round(fp32["substeps_per_s"], 2) if fp32 else None,
        "fp32_pcg_iters_per_solve": round(fp32["pcg_iters_per_solve"], 3) if fp32 else None,
        "mixed_over_fp32": round(mixed["substeps_per_s"] / fp32["substeps_per_s"], 4) if fp32 else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
