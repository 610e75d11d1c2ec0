"""The adjoint workload on the HIP path: the funnel run forward with its frames recorded, the
loss (water fraction of the bottom-right box - 0.5)^2, and its gradient with respect to the
faucet jet velocity and the funnel's 32 wall heights (tq_tape_begin / tq_backward).

The walls are built on the device from the funnel's strokes (tq_set_strokes); the wall heights
act through the SDF (the walls' push-out of particles, cena.js:102-117), and the SDF gradient
goes back to them through tq_strokes_vjp; the cell types the SDF gives are held with the other
discrete decisions of the recorded run."""
import time

import numpy as np

from . import Cena, Semente, SimulacaoHIP
from .scenes import FUNNEL_SEGMENTS, passo_frame


def run_funnel(work, checkpoint_every=10, tol=1e-6, solve_tol=1e-6, maxit=200, keep_emitted=False, grad=True,
               sim_options=None, jets=None, schedule=None, tape_from=0):
    """Runs the funnel workload (scenes.funnel) forward and, with grad, backward.  Returns a dict:
    loss, frac, g_vjet (2), g_heights (32), g_sdf, the sim (left in its end state, tape open),
    the substep schedule, and with keep_emitted the emitted particles' positions per frame.

    jets: per-frame jet velocities (vx, vy) in place of the workload's one jet; schedule: the
    substep count of every frame (default: passoMaximo from the running top speed, as the app);
    tape_from: the frame at which recording starts, so the gradient is that of the loss with
    respect to the jets of frames tape_from.. and the state at that frame (a truncated window;
    g_vjet sums the recorded calls, g_faucet_calls has one row per recorded call)."""
    nx, ny, h = work.nx, work.ny, work.h
    opts = dict(iteracoes=maxit, tol=solve_tol)
    opts.update(sim_options or {})
    sim = SimulacaoHIP(Cena(nx, ny, h), **opts)
    sim.definir_tracos(work.strokes)  # the walls built on the device
    if len(work.particles):
        sim.adicionar(np.asarray(work.particles, np.float32), False)
    f = work.extra["faucet"]
    sem = Semente(f["seed"])
    H = ny * h
    fixed = schedule
    schedule, emitted = [], []
    t0 = time.perf_counter()
    for k in range(work.frames):
        if grad and k == tape_from:
            sim.tape_begin(checkpoint_every)
        if k < f["frames"]:
            pvx, pvy = f["pvx"], f["pvy"]
            if jets is not None:  # principal.js:302: jet = 0.5 pointer + (0, -0.7)
                pvx, pvy = 2 * jets[k][0], 2 * (jets[k][1] + 0.7)
            n0 = sim.n
            sim.emitirTorneira(f["x"], f["y"], pvx, pvy, f["raio"], f["absorcao"], work.dt_frame, sem)
            if keep_emitted and k >= tape_from:
                emitted.append(sim.particulas()[n0:, :2].astype(np.float64))
        elif keep_emitted and k >= tape_from:
            emitted.append(np.zeros((0, 2)))
        nsub = fixed[k] if fixed is not None else passo_frame(h, H, sim.vMax, work.dt_frame)
        schedule.append(nsub)
        sim.avancar(work.dt_frame, nsub)
    sim.sync()
    t_forward = time.perf_counter() - t0
    box = work.extra["box"]
    target = work.extra["target"]
    out = {"sim": sim, "schedule": schedule, "emitted": emitted, "sdf": sim.ler_sdf(), "t_forward": t_forward}
    if not grad:
        frac = sim.water_fraction(box)
        out.update(loss=(frac - target) ** 2, frac=frac)
        return out
    frac, g = sim.water_fraction(box, grad=True)
    loss = (frac - target) ** 2
    t0 = time.perf_counter()
    res = sim.backward(2 * (frac - target) * g, tol=tol)
    out["t_backward"] = time.perf_counter() - t0
    g_sdf = res["sdf"].astype(np.float64)
    # a wall height moves both ends of its stroke: d/dy0 + d/dy1 (chain rule on the device)
    gs = sim.tracos_vjp(work.strokes, res["sdf"])
    g_heights = (gs[:, 1] + gs[:, 3])[: 2 * FUNNEL_SEGMENTS]
    out.update(loss=loss, frac=frac, g_vjet=res["faucet"].sum(axis=0), g_faucet_calls=res["faucet"],
               g_heights=g_heights, g_sdf=g_sdf, g_state0=res["state0"])
    return out
