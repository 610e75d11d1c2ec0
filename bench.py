#!/usr/bin/env python3
"""Benchmark of the FLIP/APIC substep on one MI355X.

Workload: the Ultra sealed hourglass (896x504 cells, 1.59 mm, ~0.75 M particles after the
wall check), frames of 1/480 s split into substeps by the reference's adaptive bound, the
pressure solved by MGPCG to a relative residual of 1e-6.  One bench step = one frame (air
update + its substeps).  The reported value is the FP32 run (FP32 PCG and multigrid, FP64 dots
and bubble unknowns), the library's default and fastest mode; the mixed-precision run (multigrid
vectors stored in FP16) replays the same substep schedule and is reported beside it with the
ratio the metric names.

Timed runs use eager launches with HIP-event timing of the dominant kernel on the library's
stream (the launch with the largest share of device time in a short untimed probe of the same
workload); the same runs are repeated with the PCG iterations replayed from captured HIP graphs
and reported beside them.  A bounded sample of the same workload on the oracle (single-threaded
C++ restatement of the reference CPU engine) is the CPU baseline.

Prints ONE JSON line.  --gpus N > 1 is not supported (one GPU per box).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tanque-de-agua_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_ultra_hourglass.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(tq, work, precision, warmup, steps, schedule=None, time_kernel=None, use_graph=False, tol=1e-6,
            maxit=200, **opts):
    from tanque_hip.scenes import passo_frame
    cena = work.cena()
    sim = tq.SimulacaoHIP(cena, precision=precision, tol=tol, iteracoes=maxit, use_graph=use_graph, **opts)
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
    if time_kernel:  # "*" times every kernel
        sim.timing_enable(True, only=None if time_kernel == "*" else time_kernel)
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
        "last_true_rel_residual": st1["true_resid_iterate"] / st1["norm_b"] if st1["norm_b"] > 0 else 0.0,
        "last_true_rel_residual_binary32_q": st1["true_resid"] / st1["norm_b"] if st1["norm_b"] > 0 else 0.0,
        "last_rel_residual": st1["norm"] / st1["norm_b"] if st1["norm_b"] > 0 else 0.0,
        "restarts_last_solve": st1["restarts"],
        "active_bubbles": st1["active_bubbles"], "levels": st1["levels"],
        "schedule": sched, "timing": timing,
    }
    sim.close()
    return res


def run_autotune(tq, work, precision, frames=8, iterations=10):
    """tq_autotune on the workload's pressure system after `frames` frames: the chosen launch
    shapes (cached by tq_create for this grid, so the runs that follow use them) and the solve
    time of the defaults and of the choice."""
    from tanque_hip.scenes import passo_frame
    cena = work.cena()
    sim = tq.SimulacaoHIP(cena, precision=precision, tol=1e-6, iteracoes=200)
    sim.adicionar(work.particles, False)
    for _ in range(frames):
        sim.avancar(work.dt_frame, passo_frame(work.h, cena.altura, sim.vMax, work.dt_frame))
    r = sim.autotune(iterations)
    sim.close()
    r["gain"] = round(r["default_us"] / r["tuned_us"], 4) if r["tuned_us"] > 0 else None
    r["iterations_per_trial_solve"] = iterations
    return r


def run_alto(tq, work, precision, tol=1e-6, maxit=200, schedule=None):
    """The Alto config end to end: each frame the faucet emits (while on) and the frame's
    substeps run, as the app does (principal.js:301-321, motor.js:98-107)."""
    from tanque_hip.scenes import passo_frame
    f = work.extra["faucet"]
    cena = work.cena()
    sim = tq.SimulacaoHIP(cena, precision=precision, tol=tol, iteracoes=maxit)
    sem = tq.Semente(f["seed"])
    sim.sync()
    st0 = sim.solver_stats()
    t0 = time.perf_counter()
    subs = 0
    psteps = 0.0
    emitted = 0
    sched = []
    for k in range(work.frames):
        if k < f["frames"]:
            n, _ = sim.emitirTorneira(f["x"], f["y"], f["pvx"], f["pvy"], f["raio"], f["absorcao"], work.dt_frame, sem)
            emitted += n
        sub = schedule[k] if schedule is not None else passo_frame(work.h, cena.altura, sim.vMax, work.dt_frame)
        sched.append(sub)
        n0 = sim.n
        sim.avancar(work.dt_frame, sub)
        subs += sub
        psteps += sub * n0
    sim.sync()
    t1 = time.perf_counter()
    st1 = sim.solver_stats()
    res = {"substeps_per_s": round(subs / (t1 - t0), 2), "particle_steps_per_s": round(psteps / (t1 - t0), 1),
           "frames": work.frames, "substeps": subs, "emitted": emitted, "particles_end": sim.n,
           "pcg_iters_per_solve": round((st1["total_iterations"] - st0["total_iterations"]) /
                                        max(1, st1["solves"] - st0["solves"]), 3)}
    sim.close()
    return res, sched


def run_batched(tq, frames=None, threads=8):
    """The batched configuration (64 maze tanks of 64-384 x 48-256 cells, fixed 3 PCG iterations as
    the reference, frames of 1/240 s) over all its frames, from fresh identical states, in three
    launch models:
      sequential  tank after tank from one thread, each tank's frame (avancar) waited for before
                  the next tank's;
      threads     every tank on its own stream, the frames enqueued from 8 host threads
                  (tq_batch_create);
      table       one table launch per kernel of the step over all 64 tanks
                  (tq_batch_create_table), issued one by one; "table_graph" replays each frame
                  from a HIP graph captured from the first frame with the same launch list.
    Substep counts come from every tank's vMax after the previous frame (passo_frame).
    tank-steps/s = tanks x frames / s.  The table run ends bit-equal to the sequential one (every
    tank's particles compared).  Then the mixed-precision preconditioner in table mode."""
    from tanque_hip import scenes
    from tanque_hip.scenes import passo_frame
    ts = scenes.batched_mazes()
    frames = frames or ts[0].frames
    res = {"tanks": len(ts), "cells": sum(t.nx * t.ny for t in ts), "frames_timed": frames,
           "of_frames": ts[0].frames, "threads": threads, "pcg": "3 iterations (reference)"}
    dt = ts[0].dt_frame

    def run(mode, prec):
        cenas = [t.cena() for t in ts]
        b = tq.Batch(cenas, [{"iteracoes": 3, "precision": prec}] * len(ts), threads=threads,
                     mode="threads" if mode in ("sequential", "threads") else "table")
        if mode == "table_graph":
            b.use_graphs(True)
        for t, s in zip(ts, b.tanks):
            s.adicionar(t.particles, False)
        res["particles"] = sum(s.n for s in b.tanks)
        table = b.mode == "table"

        def frame():
            vm = b.vmax() if table else [s.vMax for s in b.tanks]
            subs = [passo_frame(t.h, c.altura, v, dt) for t, c, v in zip(ts, cenas, vm)]
            if mode == "sequential":
                for s, k in zip(b.tanks, subs):
                    s.avancar(dt, k)
            else:
                b.step(dt, subs)
            return sum(subs)

        if table:
            b.stats(reset=True)
        t0 = time.perf_counter()
        subs = sum(frame() for _ in range(frames))
        el = time.perf_counter() - t0
        out = {"tank_steps_per_s": round(len(ts) * frames / el, 1), "tank_substeps_per_s": round(subs / el, 1),
               "seconds": round(el, 3), "mean_substeps": round(subs / frames / len(ts), 3)}
        if table:
            st = b.stats()
            out["host_ms_per_frame"] = {k[:-2]: round(st[k] / frames * 1e3, 3)
                                        for k in ("record_s", "plan_s", "stage_s", "issue_s", "wait_s")}
            out["table_launches_last_frame"] = st["launches"]
            out["record_bytes_last_frame"] = int(st["upload_bytes"])
            out["frame_graphs"] = {"captured": st["graph_builds"], "replayed": st["graph_replays"]}
        return b, out

    seq, res["sequential"] = run("sequential", tq.PREC_FP32)
    seq_parts = [s.particulas().view(np.uint32).copy() for s in seq.tanks]
    seq.close()
    for mode in ("threads", "table_graph", "table"):
        b, res[mode] = run(mode, tq.PREC_FP32)
        if mode == "table":
            res["table_bit_equal_to_sequential"] = all(
                np.array_equal(s.particulas().view(np.uint32), p) for s, p in zip(b.tanks, seq_parts))
        b.close()
        log("batched %s: %s" % (mode, res[mode]))
    res["batched"] = res["table"]
    res["batched_over_sequential"] = round(res["table"]["tank_steps_per_s"] / res["sequential"]["tank_steps_per_s"], 3)
    res["threads_over_sequential"] = round(res["threads"]["tank_steps_per_s"] / res["sequential"]["tank_steps_per_s"], 3)
    b, res["mixed_table"] = run("table", tq.PREC_MIXED)
    b.close()
    res["mixed_batched_over_fp32_batched"] = round(res["mixed_table"]["tank_steps_per_s"] /
                                                   res["table"]["tank_steps_per_s"], 3)
    return res


def run_adjoint(tq, fd_h=1e-3, checkpoint_every=10, fd=True):
    """The adjoint configuration (funnel 256x144, 120 frames of 1/240 s, faucet jet + 32 wall
    heights, loss (box water fraction - 0.5)^2): the forward alone, the forward recorded with
    checkpoints every 10 frames, the backward (recompute between checkpoints + reverse kernels),
    and central differences of the plain forward at step fd_h for the two jet components and the
    heights of the two wall segments at the spout."""
    from tanque_hip import scenes
    from tanque_hip.adjunto import run_funnel
    w = scenes.funnel()
    plain = run_funnel(w, grad=False)
    subs = int(sum(plain["schedule"]))
    plain["sim"].close()
    out = run_funnel(w, checkpoint_every=checkpoint_every)
    n_end = out["sim"].n
    info = out["sim"].tape_info()
    out["sim"].close()
    res = {"workload": w.description + ", loss box %s, target 0.5" % (w.extra["box"],),
           "frames": w.frames, "substeps": subs, "particles_end": n_end, "checkpoints": info["checkpoints"],
           "forward_s": round(plain["t_forward"], 4), "forward_substeps_per_s": round(subs / plain["t_forward"], 1),
           "recorded_forward_s": round(out["t_forward"], 4), "backward_s": round(out["t_backward"], 4),
           "backward_over_forward": round(out["t_backward"] / plain["t_forward"], 3),
           "loss": out["loss"], "frac": out["frac"],
           "grad_jet": [float(v) for v in out["g_vjet"]],
           "grad_heights_norm": float(np.linalg.norm(out["g_heights"])),
           "grad_heights_first4": [float(v) for v in out["g_heights"][:4]]}
    if not fd:
        return res
    fd = {}
    v0 = np.array([0.5 * w.extra["faucet"]["pvx"], 0.5 * w.extra["faucet"]["pvy"] - 0.7])
    for k, name in ((0, "jet_vx"), (1, "jet_vy")):
        L = []
        for sg in (1, -1):
            v = v0.copy()
            v[k] += sg * fd_h
            r = run_funnel(scenes.funnel(vjet=tuple(v)), grad=False)
            r["sim"].close()
            L.append(r["loss"])
        fd[name] = {"fd": (L[0] - L[1]) / (2 * fd_h), "adjoint": float(out["g_vjet"][k])}
    for k in (15, 31):
        L = []
        for sg in (1, -1):
            th = np.zeros(2 * scenes.FUNNEL_SEGMENTS)
            th[k] = sg * fd_h
            r = run_funnel(scenes.funnel(heights=th), grad=False)
            r["sim"].close()
            L.append(r["loss"])
        fd["height_%d" % k] = {"fd": (L[0] - L[1]) / (2 * fd_h), "adjoint": float(out["g_heights"][k])}
    res["central_differences"] = {"h": fd_h, "of": "the plain forward (every decision free), whole run, configured box",
                                  "values": fd}
    res["window"] = adjoint_window(tq, fd_h=fd_h)
    return res


FUNNEL_BOX = (100, 55, 240, 120)  # the funnel itself: the jet's water sits there within 20 frames


def adjoint_window(tq, W=5, fd_h=1e-3, box=FUNNEL_BOX, descent_steps=4, step=0.05):
    """A truncated-window objective the adjoint differentiates: the same 120-frame funnel run with
    the substep schedule of the base run, the jet of the last W frames a free parameter and the
    loss (water fraction of `box` - 0.5)^2 at the end.  Its gradient is the sum of the recorded
    faucet calls' gradients of a backward over the window (tape from frame 120 - W).  Checked
    against central differences of the forward at step fd_h, then a few normalised gradient
    steps of `step` m/s on the window's jet, each with the loss change the gradient predicted."""
    from tanque_hip import scenes
    from tanque_hip.adjunto import run_funnel
    w = scenes.funnel()
    w.extra["box"] = box
    base = run_funnel(w, grad=False)
    sched = base["schedule"]
    base["sim"].close()
    v0 = np.array([0.5 * w.extra["faucet"]["pvx"], 0.5 * w.extra["faucet"]["pvy"] - 0.7])
    F0 = w.frames - W

    def evaluate(vw, grad):
        jets = [tuple(v0)] * F0 + [tuple(vw)] * W
        r = run_funnel(w, checkpoint_every=0, grad=grad, jets=jets, schedule=sched, tape_from=F0)
        r["sim"].close()
        return r

    r0 = evaluate(v0, True)
    g = np.asarray(r0["g_vjet"], np.float64)
    fd = []
    for k in range(2):
        L = []
        for sg in (1, -1):
            v = v0.copy()
            v[k] += sg * fd_h
            L.append(evaluate(v, False)["loss"])
        fd.append((L[0] - L[1]) / (2 * fd_h))
    fd = np.array(fd)
    err = float(np.abs(fd - g).max() / max(np.abs(g).max(), 1e-300))
    steps = []
    v, L, gv = v0.copy(), r0["loss"], g
    for _ in range(descent_steps):
        d = -gv / max(np.linalg.norm(gv), 1e-300) * step
        pred = float(gv @ d)
        v = v + d
        r = evaluate(v, True)
        steps.append({"jet": [round(float(x), 5) for x in v], "loss": r["loss"], "predicted_change": pred,
                      "actual_change": r["loss"] - L})
        L, gv = r["loss"], np.asarray(r["g_vjet"], np.float64)
    return {"objective": "loss at frame %d with the jet of frames %d..%d free (box %s, target 0.5, base substep "
                         "schedule)" % (w.frames, F0, w.frames - 1, box),
            "loss": r0["loss"], "frac": r0["frac"], "adjoint": [float(x) for x in g],
            "fd": [float(x) for x in fd], "fd_h": fd_h, "max_abs_err_over_max_component": err,
            "descent": {"step_m_per_s": step, "loss_start": r0["loss"], "steps": steps}}


def run_size_sweep(tq, sizes=((896, 504), (1792, 1008), (4096, 2304)), warmup=2, frames=3, autotune=True):
    """The hourglass at growing sizes (past the 256 MB last-level cache at 4096x2304): substeps/s
    in FP32 and in the mixed mode on the same schedule, and, from a run with every kernel timed by
    HIP events on the library stream, the five kernels with the most time and their algorithmic
    bandwidth against the 8 TB/s peak."""
    from tanque_hip import scenes
    out = []
    for nx, ny in sizes:
        w = scenes.ultra_hourglass() if (nx, ny) == (896, 504) else scenes.scaled_hourglass(nx, ny)
        tune = run_autotune(tq, w, tq.PREC_FP32, frames=2) if autotune else None
        r32 = run_gpu(tq, w, tq.PREC_FP32, warmup, frames)
        timed = run_gpu(tq, w, tq.PREC_FP32, warmup, frames, schedule=r32["schedule"], time_kernel="*")
        rmx = run_gpu(tq, w, tq.PREC_MIXED, warmup, frames, schedule=r32["schedule"])
        row = {"grid": "%dx%d" % (nx, ny), "cells": nx * ny, "particles": r32["particles_start"],
               "substeps": r32["substeps"], "fp32_substeps_per_s": round(r32["substeps_per_s"], 2),
               "fp32_pcg_iters_per_solve": round(r32["pcg_iters_per_solve"], 2),
               "mixed_substeps_per_s": round(rmx["substeps_per_s"], 2),
               "mixed_pcg_iters_per_solve": round(rmx["pcg_iters_per_solve"], 2), "autotune": tune}
        tim = timed["timing"]
        total = sum(v["ms"] for v in tim.values())
        top = []
        for name, v in sorted(tim.items(), key=lambda kv: -kv[1]["ms"])[:5]:
            avg = v["ms"] / max(v["launches"], 1) * 1e-3
            gbs = v["bytes_per_launch"] / avg / 1e9 if avg > 0 else 0.0
            top.append({"kernel": name, "share": round(v["ms"] / total, 3), "us": round(avg * 1e6, 2),
                        "GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
        row["timed_kernel_ms_per_substep"] = round(total / max(timed["substeps"], 1), 3)
        row["top_kernels"] = top
        row["kernels_us_GBs"] = {name: [round(v["ms"] / max(v["launches"], 1) * 1e3, 2),
                                        round(v["bytes_per_launch"] / (v["ms"] / max(v["launches"], 1) * 1e-3) / 1e9, 1)
                                        if v["ms"] > 0 else 0.0] for name, v in sorted(tim.items())}
        log("size sweep %s" % row)
        out.append(row)
    return out


def dominant_kernel(tq, work, precision, frames=3):
    """Name of the kernel family with the largest device time on a few frames of the workload."""
    from tanque_hip.scenes import passo_frame
    cena = work.cena()
    sim = tq.SimulacaoHIP(cena, precision=precision, tol=1e-6, iteracoes=200)
    sim.adicionar(work.particles, False)
    sim.avancar(work.dt_frame, passo_frame(work.h, cena.altura, sim.vMax, work.dt_frame))
    sim.sync()
    sim.timing_enable(True)
    for _ in range(frames):
        sim.avancar(work.dt_frame, passo_frame(work.h, cena.altura, sim.vMax, work.dt_frame))
    sim.sync()
    t = sim.timing_read()
    sim.close()
    name = max(t, key=lambda k: t[k]["ms"])
    return name, t


def host_cpu_info():
    """nproc and the lscpu summary of the host the CPU baseline ran on."""
    import subprocess
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keep = ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)", "CPU max MHz",
                "L3 cache")
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keep:
                info["lscpu " + k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        info["lscpu"] = "unavailable"
    return info


def cpu_baseline(work, budget_s=15.0, tol=1e-6, maxit=200):
    """The oracle (single-threaded C++ restatement of the reference CPU engine) on a bounded
    sample of the same workload: the first frames of the same scene, same particles.  Driven from
    Node, the reference's host (oracle/bench_oracle.js over the oracle's Node-API addon), as a
    child process; from Python where node or the addon is missing."""
    import shutil
    import subprocess
    import tempfile
    from oracle import oracle as orc
    orc.build()
    here = os.path.dirname(os.path.abspath(__file__))
    harness = os.path.join(here, "oracle", "bench_oracle.js")
    node = shutil.which("node")
    if node and os.path.exists(orc.NODE_ADDON):
        with tempfile.TemporaryDirectory() as d:
            pf = os.path.join(d, "particles.f32")
            np.ascontiguousarray(work.particles, np.float32).tofile(pf)
            wf = os.path.join(d, "workload.json")
            with open(wf, "w") as f:
                json.dump({"nx": work.nx, "ny": work.ny, "h": work.h,
                           "strokes": [[float(v) for v in s[:5]] + [bool(s[5]) if len(s) > 5 else False]
                                       for s in work.strokes],
                           "particles_file": pf, "dt_frame": work.dt_frame, "frames": work.frames,
                           "budget_s": budget_s, "tol": tol, "maxit": maxit}, f)
            r = subprocess.run([node, harness, wf], capture_output=True, text=True, timeout=budget_s + 300)
        if r.returncode != 0:
            raise RuntimeError("bench_oracle.js failed: %s" % r.stderr[-2000:])
        out = json.loads(r.stdout.strip().splitlines()[-1])
        frames, subs, el = out["frames"], out["substeps"], out["seconds"]
        driver = "node %s (oracle/bench_oracle.js over oracle/oracle_node.node)" % (
            subprocess.run([node, "--version"], capture_output=True, text=True).stdout.strip())
    else:
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
        driver = "python (ctypes over oracle/liboracle.so)"
    return {"value": round(subs / el, 4), "unit": "substeps/s", "cores": 1, "kind": "port", "driver": driver,
            "host": host_cpu_info(),
            "sample": "first %d frames (%d substeps) of the same workload on the oracle (JS-exact binary64 "
                      "compute / binary32 storage, PCG to 1e-6), %.1f s on one host core" % (frames, subs, el)}


def copy_calibration(nbytes=256 << 20):
    """One device-to-device copy of known size (the runtime's copy kernel), so that a
    rocprofv3 --pmc run of this command can scale FETCH_SIZE / WRITE_SIZE to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 0, ctypes.c_size_t(nbytes))
    if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log("copy calibration: %d bytes device to device" % nbytes)


def profile_for(name):
    """The committed rocprofv3 figures of one timing name (profiles/traffic_ultra_hourglass.json):
    HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, calibrated on a copy of known size) and the
    kernel-trace average duration of the same launches."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    return d.get("kernels", {}).get(name)


# ------------------------------------------------------------------ rocprofv3 summaries (profiles/)
def _csv_rows(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _find(d, name):
    """The rocprofv3 CSV of `name` (kernel_trace, counter_collection) under directory d."""
    for root, _, files in os.walk(d):
        for fn in files:
            if fn.endswith(name + ".csv"):
                return os.path.join(root, fn)
    raise FileNotFoundError("%s: no *%s.csv" % (d, name))


def _grid_total(r):
    if "Grid_Size" in r:
        return int(r["Grid_Size"])
    return int(r.get("Grid_Size_X", 1)) * int(r.get("Grid_Size_Y", 1)) * int(r.get("Grid_Size_Z", 1))


def body_name(mangled):
    """The body of a launch wrapper (tq::k_one / k_one_w launch one tank, tq::k_tab / k_tab_w a
    tank table) from its mangled symbol, which the profiler leaves mangled (the wrappers take the
    body as an auto template argument): "tq::k_bra<float, float, float, float, true>", with
    " [table]" for a table launch; None for other symbols.  The bodies' template arguments are
    float, __half (its later uses are substitutions S<n>_), bool and int literals."""
    import re
    m = re.match(r"_ZN2tq\d+k_(one|tab)(?:_w)?I.*?XadL_ZNS_(\d+)", mangled)
    if not m:
        return None
    pos, ln = m.end(), int(m.group(2))
    body, rest = mangled[pos:pos + ln], mangled[pos + ln:]
    args = []
    if rest.startswith("I"):
        i = 1
        while i < len(rest) and rest[i] != "E":
            if rest[i] == "f":
                args.append("float")
                i += 1
            elif rest.startswith("6__half", i):
                args.append("__half")
                i += 7
            elif rest[i] == "S":
                args.append("__half")
                i = rest.index("_", i) + 1
            elif rest.startswith("Lb", i):
                args.append("true" if rest[i + 2] == "1" else "false")
                i = rest.index("E", i) + 1
            elif rest.startswith("Li", i):
                j = rest.index("E", i)
                args.append(rest[i + 2:j])
                i = j + 1
            else:
                args.append("?")
                break
    name = "tq::%s<%s>" % (body, ", ".join(args)) if args else "tq::" + body
    return name + (" [table]" if m.group(1) == "tab" else "")


def _kname(r):
    n = r["Kernel_Name"].split("(")[0].strip()
    b = body_name(n)
    if b:
        return b
    return n[5:] if n.startswith("void ") else n


def solver_kernel_map(nx, ny):
    """Timing name -> (kernel symbol, total grid in threads) of the FP32 solver kernels as the
    library launches them on an nx x ny grid with the default launch shapes (levels halve while
    min(nx, ny) > 12; the fused down kernel has 9 waves per block, the others 256 threads)."""
    dims = [(nx, ny)]
    while min(dims[-1]) > 12:
        dims.append(((dims[-1][0] + 1) // 2, (dims[-1][1] + 1) // 2))
    t16 = lambda n: (n + 15) // 16
    groups = lambda a, b, d: ((a + d + 15) // 16 + 1) // 2 * (((b + d + 15) // 16 + 1) // 2)
    m = {"pcg_spmv": ("tq::k_spmv", min(((nx + 63) // 64) * ((ny + 15) // 16), 1024) * 256),
         "pcg_update_mg_in_l0": ("tq::k_in_a0<float, true, true>", min(groups(nx, ny, 0), 1024) * 256),
         "mg_pass_z_l0": ("tq::k_pass_z<float>", groups(nx, ny, 0) * 256),
         "mg_prolong_c_l0": ("tq::k_prolong_pass<float, float, float>", groups(nx, ny, 8) * 256)}
    for l in range(len(dims) - 1):
        a, b = dims[l + 1]
        for passa in ("true", "false"):
            m.setdefault("mg_down_l%d" % l, None)
            key = ("tq::k_bra<float, float, float, float, %s>" % passa, t16(a) * t16(b) * 576)
            m["mg_down_l%d:%s" % (l, passa)] = key
        if l > 0:
            m["mg_up_l%d" % l] = ("tq::k_pcd<float, float, float, false>", t16(dims[l][0]) * t16(dims[l][1]) * 256)
    return {k: v for k, v in m.items() if v is not None}


def summarize_profiles(fetch_dir, write_dir, trace_dir, nx, ny, workload, commands, copy_bytes=256 << 20):
    """The traffic (FETCH_SIZE, WRITE_SIZE in KiB per dispatch, scaled so a device copy of
    copy_bytes reads and writes copy_bytes) and kernel-trace durations of every (kernel symbol,
    grid) of a run, with the solver kernels also under their timing names."""
    import statistics
    def counters(d, counter):
        per = {}
        calib = None
        for r in _csv_rows(_find(d, "counter_collection")):
            if r.get("Counter_Name") != counter:
                continue
            v = float(r["Counter_Value"])
            k = (_kname(r), _grid_total(r))
            if "copyBuffer" in k[0] and (calib is None or v > calib):
                calib = v
            per.setdefault(k, []).append(v)
        return per, calib
    fetch, fcal = counters(fetch_dir, "FETCH_SIZE")
    write, wcal = counters(write_dir, "WRITE_SIZE")
    ffac = copy_bytes / (fcal * 1024.0) if fcal else 1.0
    wfac = copy_bytes / (wcal * 1024.0) if wcal else 1.0
    dur = {}
    for r in _csv_rows(_find(trace_dir, "kernel_trace")):
        k = (_kname(r), _grid_total(r))
        dur.setdefault(k, []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    table = []
    for k, ds in dur.items():
        f = fetch.get(k, [])
        w = write.get(k, [])
        table.append({"symbol": k[0], "grid_total": k[1], "dispatches": len(ds),
                      "total_ms": round(sum(ds) / 1e3, 3), "avg_us": round(sum(ds) / len(ds), 3),
                      "median_us": round(statistics.median(ds), 3),
                      "fetch_bytes_per_launch": round(sum(f) / len(f) * 1024 * ffac) if f else None,
                      "write_bytes_per_launch": round(sum(w) / len(w) * 1024 * wfac) if w else None})
    table.sort(key=lambda e: -e["total_ms"])
    return {"workload": workload, "commands": commands,
            "calibration": {"copy_bytes": copy_bytes, "fetch_kib": fcal, "write_kib": wcal,
                            "fetch_factor": ffac, "write_factor": wfac},
            "kernels": named_kernels(table, nx, ny), "by_symbol_grid": table}


def named_kernels(table, nx, ny):
    """The solver kernels of a (symbol, grid) table under their timing names."""
    byk = {(e["symbol"][5:] if e["symbol"].startswith("void ") else e["symbol"], e["grid_total"]): e for e in table}
    kernels = {}
    for name, key in solver_kernel_map(nx, ny).items():
        e = byk.get(key)
        if not e:
            continue
        base = name.split(":")[0]
        if base in kernels:
            continue
        hb = (e["fetch_bytes_per_launch"] or 0) + (e["write_bytes_per_launch"] or 0)
        kernels[base] = {"symbol": key[0], "grid_total": key[1], "dispatches": e["dispatches"],
                         "fetch_bytes_per_launch": e["fetch_bytes_per_launch"],
                         "write_bytes_per_launch": e["write_bytes_per_launch"],
                         "hbm_bytes_per_launch": hb or None, "rocprof_avg_us": e["avg_us"],
                         "rocprof_median_us": e["median_us"]}
    return kernels


def run_batched_table_only(tq, frames):
    """The table launch model of the batched configuration alone, for a profiler run: the
    configured 64 mazes stepped `frames` frames from fresh states after one untimed frame."""
    from tanque_hip import scenes
    from tanque_hip.scenes import passo_frame
    ts = scenes.batched_mazes()
    cenas = [t.cena() for t in ts]
    b = tq.Batch(cenas, [{"iteracoes": 3}] * len(ts), threads=8, mode="table")
    for t, s in zip(ts, b.tanks):
        s.adicionar(t.particles, False)
    dt = ts[0].dt_frame
    subs = 0
    t0 = 0.0
    for f in range(frames + 1):
        if f == 1:
            t0 = time.perf_counter()
        k = [passo_frame(t.h, c.altura, v, dt) for t, c, v in zip(ts, cenas, b.vmax())]
        b.step(dt, k)
        subs += sum(k) if f else 0
    el = time.perf_counter() - t0
    b.close()
    return {"frames": frames, "tank_steps_per_s": round(len(ts) * frames / el, 1), "substeps": subs}


def summarize_batch_trace(trace_dir, frames, commands):
    """Markdown table of the kernels of a table-batch run's rocprofv3 kernel trace (the run of
    run_batched_table_only with frames + 1 frames): per body, launches and device time per frame."""
    import statistics
    rows = _csv_rows(_find(trace_dir, "kernel_trace"))
    per = {}
    for r in rows:
        k = _kname(r)
        per.setdefault(k, []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    fr = frames + 1
    tab = {k: v for k, v in per.items() if k.endswith("[table]")}
    total = sum(sum(v) for v in tab.values()) / fr
    lines = ["# Kernel time of the table-batched configuration (64 maze tanks, 2.39 M cells)", "",
             "Commands (from the repository root, on one MI355X):", "", "```"] + commands + ["```", "",
             "%d frames (the untimed first frame included); table launches only (the set-up's per-tank "
             "launches are left out).  Device time of the table kernels: %.1f us per frame." % (fr, total), "",
             "| body | launches per frame | avg us | median us | us per frame | share |",
             "|---|---:|---:|---:|---:|---:|"]
    for k, v in sorted(tab.items(), key=lambda kv: -sum(kv[1])):
        s = sum(v) / fr
        lines.append("| `%s` | %.1f | %.2f | %.2f | %.1f | %.1f %% |" % (
            k.replace(" [table]", ""), len(v) / fr, sum(v) / len(v), statistics.median(v), s, 100 * s / total))
    return "\n".join(lines) + "\n"


def summarize_trace(trace_dir, title, commands, note, top=24):
    """Markdown table of every kernel of a rocprofv3 --kernel-trace run (by body for the library's
    launch wrappers): calls, total and average device time, share; the adjoint's reverse kernels
    (bodies k_adj_* and k_ten_*) listed in full below it."""
    rows = _csv_rows(_find(trace_dir, "kernel_trace"))
    per = {}
    for r in rows:
        per.setdefault(_kname(r), []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    total = sum(sum(v) for v in per.values())
    order = sorted(per.items(), key=lambda kv: -sum(kv[1]))

    def table(items):
        out = ["| kernel | calls | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
        for k, v in items:
            out.append("| `%s` | %d | %.2f | %.2f | %.1f |" % (k, len(v), sum(v) / 1e3, sum(v) / len(v),
                                                             100 * sum(v) / total))
        return out
    lines = ["# " + title, "", "Command (from the repository root, on one MI355X):", "", "```"] + commands + \
        ["```", "", note, "", "Device time of all kernels: %.1f ms." % (total / 1e3), ""] + table(order[:top])
    rev = [(k, v) for k, v in order if k.startswith(("tq::k_adj_", "tq::k_ten_"))]
    if rev:
        lines += ["", "The reverse kernels (all of them):", ""] + table(rev)
    return "\n".join(lines) + "\n"


WAVE_COUNTERS = ("SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY",
                 "SQ_ACTIVE_INST_ANY", "GRBM_GUI_ACTIVE")


def summarize_wave_counters(runs, commands, note, top=16):
    """Markdown tables of a rocprofv3 --pmc run of WAVE_COUNTERS per (kernel symbol, grid): waves
    per launch, and the wave-cycles split into cycles issuing (SQ_ACTIVE_INST_ANY) and waiting
    (SQ_WAIT_ANY: for data or a dependency; SQ_WAIT_INST_ANY: for an issue slot), each over
    SQ_WAVE_CYCLES; SQ_WAVE_CYCLES over SQ_BUSY_CYCLES is the mean count of resident waves (in the
    counters' own units, comparable between kernels and runs).  `runs` is a list of
    (counter directory, title)."""
    lines = ["# Wave residency and stall cycles of the solver kernels", "",
             "Commands (from the repository root, on one MI355X):", "", "```"] + commands + ["```", "", note, ""]
    for d, title in runs:
        per, disp, dur = {}, {}, {}
        for r in _csv_rows(_find(d, "counter_collection")):
            k = (_kname(r), _grid_total(r))
            c = per.setdefault(k, {})
            c[r["Counter_Name"]] = c.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
            disp.setdefault(k, set()).add(r["Dispatch_Id"])
            dur.setdefault(k, {})[r["Dispatch_Id"]] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        rows = []
        for k, c in per.items():
            n = len(disp[k])
            us = sum(dur[k].values()) / n
            wc = c.get("SQ_WAVE_CYCLES", 0.0) or 1.0
            rows.append((us * n, k, n, us, c.get("SQ_WAVES", 0.0) / n, c.get("SQ_ACTIVE_INST_ANY", 0.0) / wc,
                         c.get("SQ_WAIT_ANY", 0.0) / wc, c.get("SQ_WAIT_INST_ANY", 0.0) / wc,
                         c.get("SQ_WAVE_CYCLES", 0.0) / (c.get("SQ_BUSY_CYCLES", 0.0) or 1.0)))
        rows.sort(key=lambda t: -t[0])
        lines += ["## " + title, "",
                  "| kernel | grid | calls | avg us | waves / launch | issuing | waiting (data, dependency) "
                  "| waiting (issue) | resident waves |",
                  "|---|---:|---:|---:|---:|---:|---:|---:|---:|"]
        for _, k, n, us, w, a, wa, wi, res in rows[:top]:
            lines.append("| `%s` | %d | %d | %.2f | %.0f | %.3f | %.3f | %.3f | %.1f |" % (k[0], k[1], n, us, w, a, wa, wi, res))
        lines.append("")
    return "\n".join(lines) + "\n"


def profile_markdown(summary, title, top=30):
    """The kernel table of a summary as markdown (profiles/*.md)."""
    lines = ["# " + title, "", "Commands (from the repository root, on one MI355X):", "", "```"]
    lines += summary["commands"] + ["```", "",
                                    "FETCH_SIZE / WRITE_SIZE scaled by a %d-byte device copy in the same runs (factors %.3f / "
                                    "%.3f).  Per (kernel, grid in threads):" % (
                                        summary["calibration"]["copy_bytes"], summary["calibration"]["fetch_factor"],
                                        summary["calibration"]["write_factor"]), "",
                                    "| kernel | grid | calls | total ms | avg us | median us | fetch MB/launch | write MB/launch |",
                                    "|---|---:|---:|---:|---:|---:|---:|---:|"]
    for e in summary["by_symbol_grid"][:top]:
        mb = lambda v: "%.3f" % (v / 1e6) if v is not None else "-"
        lines.append("| `%s` | %d | %d | %.2f | %.2f | %.2f | %s | %s |" % (
            e["symbol"], e["grid_total"], e["dispatches"], e["total_ms"], e["avg_us"], e["median_us"],
            mb(e["fetch_bytes_per_launch"]), mb(e["write_bytes_per_launch"])))
    return "\n".join(lines) + "\n"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120, help="timed frames")
    ap.add_argument("--warmup", type=int, default=10, help="untimed frames")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision runs")
    ap.add_argument("--no-graph", action="store_true", help="skip the graph-replay runs")
    ap.add_argument("--no-variants", action="store_true", help="skip the partial-FP16 variants")
    ap.add_argument("--no-alto", action="store_true", help="skip the Alto (cup + faucet) runs")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched (64 tanks) runs")
    ap.add_argument("--batched-frames", type=int, default=0, help="frames of the batched runs (0: all 500)")
    ap.add_argument("--batched-table-only", action="store_true",
                    help="only the table-batched configuration for --batched-frames frames (default 20; for profiling)")
    ap.add_argument("--summarize-trace", nargs=3, metavar=("TRACE_DIR", "OUT_MD", "TITLE"),
                    help="write the kernel table of a rocprofv3 --kernel-trace run (with --profile-commands "
                         "and --profile-note)")
    ap.add_argument("--profile-note", default="", help="a paragraph for --summarize-trace")
    ap.add_argument("--summarize-wave-counters", nargs="+", metavar="ARG",
                    help="OUT_MD then pairs of (counter directory, title) of rocprofv3 --pmc runs of the wave "
                         "counters (with --profile-commands and --profile-note)")
    ap.add_argument("--summarize-batch-trace", nargs=2, metavar=("TRACE_DIR", "OUT_MD"),
                    help="write the kernel table of a --batched-table-only rocprofv3 --kernel-trace run")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint (funnel) runs")
    ap.add_argument("--no-autotune", action="store_true", help="run with the built-in launch shapes")
    ap.add_argument("--summarize-profiles", nargs=5, metavar=("FETCH_DIR", "WRITE_DIR", "TRACE_DIR", "OUT_JSON", "OUT_MD"),
                    help="write the traffic/duration summary of three rocprofv3 runs of this workload (--nx/--ny)")
    ap.add_argument("--profile-commands", action="append", default=[], help="the commands of the summarized runs")
    ap.add_argument("--size-sweep", action="store_true",
                    help="only the hourglass at 896x504, 1792x1008 and 4096x2304: FP32 and mixed on one schedule, "
                         "and the five kernels with the most time (every kernel timed by HIP events); prints its JSON")
    ap.add_argument("--sweep-grid", action="append", default=None, metavar="WxH",
                    help="with --size-sweep: only these grids (repeatable), e.g. 4096x2304")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the adjoint configuration's forward and backward (for kernel profiles); prints its JSON")
    ap.add_argument("--adjoint-check", action="store_true",
                    help="only the adjoint configuration with its central differences and windowed check")
    ap.add_argument("--sync-probe", action="store_true",
                    help="only the kernel-boundary / grid-barrier cost probe (tanque-de-agua_amd/lib/bench_sync)")
    ap.add_argument("--stream-probe", action="store_true",
                    help="only the HBM streaming-bandwidth probe (tanque-de-agua_amd/lib/bench_stream)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="start with a 256 MiB device copy (reference for rocprofv3 --pmc byte counters)")
    ap.add_argument("--nx", type=int, default=896)
    ap.add_argument("--ny", type=int, default=504)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import tanque_hip as tq
    from tanque_hip import scenes
    tq.load()
    for flag, probe in ((args.sync_probe, "bench_sync"), (args.stream_probe, "bench_stream")):
        if not flag:
            continue
        import subprocess
        exe = os.path.join(ROOT, "tanque-de-agua_amd", "lib", probe)
        if not os.path.exists(exe):
            raise SystemExit("%s is not built (__graft_entry__.build())" % probe)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            raise SystemExit("%s failed: %s" % (probe, r.stderr))
        print(json.dumps({probe[len("bench_"):] + "_probe": json.loads(r.stdout)}), flush=True)
        return
    if args.summarize_profiles:
        fd, wd, td, oj, om = args.summarize_profiles
        name = "hourglass %dx%d" % (args.nx, args.ny)
        sm = summarize_profiles(fd, wd, td, args.nx, args.ny, name, args.profile_commands)
        with open(oj, "w") as f:
            json.dump(sm, f, indent=1)
        with open(om, "w") as f:
            f.write(profile_markdown(sm, "Kernel time and HBM traffic, %s" % name))
        print(json.dumps({"summary": oj, "kernels": sm["kernels"]}), flush=True)
        return
    if args.summarize_wave_counters:
        om, rest = args.summarize_wave_counters[0], args.summarize_wave_counters[1:]
        if len(rest) < 2 or len(rest) % 2:
            ap.error("--summarize-wave-counters needs OUT_MD and pairs of (directory, title)")
        with open(om, "w") as f:
            f.write(summarize_wave_counters(list(zip(rest[0::2], rest[1::2])), args.profile_commands,
                                            args.profile_note))
        print(json.dumps({"summary": om}), flush=True)
        return
    if args.summarize_trace:
        td, om, title = args.summarize_trace
        with open(om, "w") as f:
            f.write(summarize_trace(td, title, args.profile_commands, args.profile_note))
        print(json.dumps({"summary": om}), flush=True)
        return
    if args.summarize_batch_trace:
        td, om = args.summarize_batch_trace
        with open(om, "w") as f:
            f.write(summarize_batch_trace(td, args.batched_frames or 20, args.profile_commands))
        print(json.dumps({"summary": om}), flush=True)
        return
    if args.batched_table_only:
        print(json.dumps({"batched_table": run_batched_table_only(tq, args.batched_frames or 20)}), flush=True)
        return
    if args.adjoint_only or args.adjoint_check:
        print(json.dumps({"adjoint": run_adjoint(tq, fd=args.adjoint_check)}), flush=True)
        return
    if args.copy_calibration:
        copy_calibration()
    if args.size_sweep:
        sizes = tuple(tuple(int(v) for v in g.lower().split("x")) for g in args.sweep_grid) if args.sweep_grid \
            else ((896, 504), (1792, 1008), (4096, 2304))
        print(json.dumps({"size_sweep": run_size_sweep(tq, sizes=sizes, autotune=not args.no_autotune)}), flush=True)
        return
    work = scenes.ultra_hourglass() if (args.nx, args.ny) == (896, 504) else scenes.scaled_hourglass(args.nx, args.ny)
    log("workload %s: %dx%d cells, %d particles seeded" % (work.name, work.nx, work.ny, len(work.particles) // 12))

    tune = None
    if not args.no_autotune:
        tune = {"fp32": run_autotune(tq, work, tq.PREC_FP32)}
        if not args.no_mixed:
            tune["mixed"] = run_autotune(tq, work, tq.PREC_MIXED)
        log("autotune: %s" % tune)
    dom, probe = dominant_kernel(tq, work, tq.PREC_FP32)
    log("dominant kernel (fp32): %s (%.1f%% of device time in the probe)" % (
        dom, 100 * probe[dom]["ms"] / sum(v["ms"] for v in probe.values())))

    fp32 = run_gpu(tq, work, tq.PREC_FP32, args.warmup, args.steps)
    log("fp32: %.1f substeps/s, %.2f PCG it/solve, last rel resid %.2e (true %.2e)" % (
        fp32["substeps_per_s"], fp32["pcg_iters_per_solve"], fp32["last_rel_residual"],
        fp32["last_true_rel_residual"]))
    sched = fp32["schedule"]
    # the same schedule again with the dominant kernel's launches bracketed by HIP events
    timed = run_gpu(tq, work, tq.PREC_FP32, args.warmup, args.steps, schedule=sched, time_kernel=dom)
    log("fp32 with %s timed: %.1f substeps/s" % (dom, timed["substeps_per_s"]))
    mixed = None
    if not args.no_mixed:
        mixed = run_gpu(tq, work, tq.PREC_MIXED, args.warmup, args.steps, schedule=sched)
        log("mixed: %.1f substeps/s, %.2f PCG it/solve, last rel resid %.2e (true %.2e)" % (
            mixed["substeps_per_s"], mixed["pcg_iters_per_solve"], mixed["last_rel_residual"],
            mixed["last_true_rel_residual"]))
    graph = {}
    if not args.no_graph:
        for lab, prec in (("fp32", tq.PREC_FP32), ("mixed", tq.PREC_MIXED)):
            if lab == "mixed" and args.no_mixed:
                continue
            r = run_gpu(tq, work, prec, args.warmup, args.steps, schedule=sched, use_graph=True)
            graph[lab] = {"substeps_per_s": round(r["substeps_per_s"], 2),
                          "pcg_iters_per_solve": round(r["pcg_iters_per_solve"], 3)}
            log("graph %s: %.1f substeps/s" % (lab, r["substeps_per_s"]))
    variants = {}
    if not args.no_variants and not args.no_mixed:
        for lab, prec, hl in (("fp16_b_only", tq.PREC_MIXED_B, -1), ("fp16_x_only_from_level2", tq.PREC_MIXED_X, 2)):
            r = run_gpu(tq, work, prec, args.warmup, args.steps, schedule=sched, mg_half_level=hl)
            variants[lab] = {"substeps_per_s": round(r["substeps_per_s"], 2),
                             "pcg_iters_per_solve": round(r["pcg_iters_per_solve"], 3)}
            log("variant %s: %.1f substeps/s, %.2f it/solve" % (lab, r["substeps_per_s"], r["pcg_iters_per_solve"]))

    kt = timed["timing"].get(dom)
    roofline = None
    if kt and kt["launches"]:
        avg_s = kt["ms"] / kt["launches"] * 1e-3
        achieved = kt["bytes_per_launch"] / avg_s / 1e9
        prof = profile_for(dom) or {}
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                    "traffic": prof.get("hbm_bytes_per_launch"), "avg_launch_us": round(avg_s * 1e6, 3),
                    "algorithmic_bytes_per_launch": kt["bytes_per_launch"], "launches": kt["launches"],
                    "rocprof_avg_us": prof.get("rocprof_avg_us"),
                    "frac_at_rocprof_duration": (round(kt["bytes_per_launch"] / (prof["rocprof_avg_us"] * 1e-6) / 1e9
                                                       / HBM_PEAK_GBS, 6) if prof.get("rocprof_avg_us") else None)}

    alto = None
    if not args.no_alto:
        wa = scenes.alto_cup()
        r32, sched_a = run_alto(tq, wa, tq.PREC_FP32)
        alto = {"workload": wa.description, "fp32": r32}
        log("alto fp32: %s" % alto["fp32"])
        if not args.no_mixed:
            alto["mixed"], _ = run_alto(tq, wa, tq.PREC_MIXED, schedule=sched_a)
            alto["mixed_over_fp32"] = round(alto["mixed"]["substeps_per_s"] / alto["fp32"]["substeps_per_s"], 4)
            log("alto mixed: %s" % alto["mixed"])

    batched = None
    if not args.no_batched:
        batched = run_batched(tq, frames=args.batched_frames or None)
        log("batched: %s" % batched)

    adjoint = None
    if not args.no_adjoint:
        adjoint = run_adjoint(tq)
        log("adjoint: %s" % adjoint)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(work, args.cpu_budget)
        log("cpu baseline: %.3f substeps/s (%s)" % (cpu["value"], cpu["sample"]))

    ms_per_step = fp32["seconds"] / args.steps * 1e3
    line = {
        "metric": "substeps/s @~450k cells, mixed vs FP32 MGPCG at 1e-6 resid; % HBM peak",
        "value": round(fp32["substeps_per_s"], 2),
        "unit": "substeps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (PCG and multigrid f32; dots, norms and bubble unknowns f64)",
        "data": "synthetic",
        "config": {"workload": "%s: %s, frames of %.6f s, PCG to rel. residual 1e-6" % (
            work.name, work.description, work.dt_frame),
            "precision": "fp32 (value); mixed reported beside it",
            "cells": work.nx * work.ny, "particles": fp32["particles_start"],
            "substeps_timed": fp32["substeps"], "pcg_iters_per_solve": round(fp32["pcg_iters_per_solve"], 3),
            "last_rel_residual": fp32["last_rel_residual"],
            "last_true_rel_residual": fp32["last_true_rel_residual"],
            "last_true_rel_residual_binary32_q": fp32["last_true_rel_residual_binary32_q"],
            "last_restarts": fp32["restarts_last_solve"],
            "mixed_last_true_rel_residual": mixed["last_true_rel_residual"] if mixed else None,
            "stop_test": ("recursive residual <= 1e-6, then the binary64 true residual of the compensated iterate "
                          "(q + q_lo) recomputed; restart from it while above 1e-6"),
            "active_bubbles_end": fp32["active_bubbles"], "mg_levels": fp32["levels"],
            "launches": "eager; a second run of the same schedule times the dominant kernel with HIP events"},
        "particle_steps_per_s": round(fp32["particle_steps_per_s"], 1),
        "timed_run_substeps_per_s": round(timed["substeps_per_s"], 2),
        "mixed_substeps_per_s": round(mixed["substeps_per_s"], 2) if mixed else None,
        "mixed_pcg_iters_per_solve": round(mixed["pcg_iters_per_solve"], 3) if mixed else None,
        "mixed_over_fp32": round(mixed["substeps_per_s"] / fp32["substeps_per_s"], 4) if mixed else None,
        "autotune": tune,
        "graph_replay": graph or None,
        "mixed_variants": variants or None,
        "alto": alto,
        "batched": batched,
        "adjoint": adjoint,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
