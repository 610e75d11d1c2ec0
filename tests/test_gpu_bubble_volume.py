"""Bubble volume over one second of the sealed hourglass (the Ultra scene at 224x126): the HIP
path in FP32 and in the mixed-precision solver against the FP64 oracle on the same substep
schedule, pressure solved to 1e-6.

The lower chamber's air starts as one bubble; the tracked volume (volumeAlvo, ar.js) and cell
count (volumeBolha) of the main bubble and the bubble count are recorded every frame.  While both
runs hold the same bubbles the main bubble's volume must stay that of the FP64 oracle's run:
relative difference of the tracked volume below 1e-4, and of the cell count within the few tens
of cells by which two runs' bubble surfaces differ.  A topology event -- a few-cell pocket split
off as air rises through the neck, or atmosphere trapped -- hands volume between bubbles
(ar.js:96-123); whether and when one happens is decided by the chaotic particle trajectories,
which differ between any two arithmetics, so the comparison runs up to the first frame at which
either run's bubble count changes (one FP32 run traps a 25-cell pocket near the end of the
second that the FP64 run does not), and that window must cover at least a quarter second.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FRAMES = 480          # 1 s at 1/480 s
RTOL_TARGET = 1e-4    # tracked volume vs FP64
RTOL_CELLS = 5e-3     # cell count vs FP64: the runs' bubble surfaces differ by a few tens of
                      # cells of ~11,000 (measured 1.9e-3 in FP32, 1.7e-3 mixed)


def bubble_volumes(sim, orc):
    """(main bubble's target, its cells, bubble count, total target, total cells of the bubbles)"""
    alvo = np.asarray(sim.read(orc.F_VALVO), np.float64)
    cells = np.asarray(sim.read(orc.F_VOLBOLHA), np.float64)
    k = int(np.argmax(alvo))
    live = alvo[2:] > 0
    return alvo[k], cells[k], int(live.sum()), alvo[2:][live].sum(), cells[2:][live].sum()


def test_bubble_volume_follows_fp64(oracle):
    import tanque_hip as tq
    from tanque_hip import scenes
    from tanque_hip.scenes import passo_frame
    orc = oracle
    w = scenes.scaled_hourglass(224, 126)
    H = w.ny * w.h
    runs = {}
    schedule = None
    for label, prec in (("fp32", tq.PREC_FP32), ("mixed", tq.PREC_MIXED)):
        sim = tq.SimulacaoHIP(w.cena(), precision=prec, tol=1e-6, iteracoes=200)
        sim.adicionar(w.particles, False)
        rec, sched = [], []
        for k in range(FRAMES):
            sub = schedule[k] if schedule else passo_frame(w.h, H, sim.vMax, w.dt_frame)
            sched.append(sub)
            sim.avancar(w.dt_frame, sub)
            rec.append(bubble_volumes(sim, orc))
        schedule = schedule or sched
        runs[label] = np.array(rec)
        sim.close()
    c = orc.Cena(w.nx, w.ny, w.h)
    for s in w.strokes:
        c.tracar(*s)
    o = orc.Simulacao(c, mode=1, iteracoes=200, tol=1e-6)
    o.adicionar(w.particles, False)
    rec = []
    for k in range(FRAMES):
        o.avancar(w.dt_frame, schedule[k])
        rec.append(bubble_volumes(o, orc))
    runs["fp64"] = np.array(rec)
    ref = runs["fp64"]
    print("\n  substeps %d; FP64 bubble: target %.1f -> %.1f cells, cells %d -> %d, bubbles %d -> %d"
          % (sum(schedule), ref[0, 0], ref[-1, 0], ref[0, 1], ref[-1, 1], ref[0, 2], ref[-1, 2]))
    for label in ("fp32", "mixed"):
        r = runs[label]
        ev = np.nonzero((r[:, 2] != r[0, 2]) | (ref[:, 2] != ref[0, 2]))[0]
        k = int(ev[0]) if len(ev) else FRAMES
        dt_ = np.abs(r[:k, 0] - ref[:k, 0]).max() / ref[:k, 0].max()
        dc = np.abs(r[:k, 1] - ref[:k, 1]).max() / ref[:k, 1].max()
        drift = (r[k - 1, 1] - r[k - 1, 0]) / r[k - 1, 0]
        after = (np.abs(r[k:, 3] - ref[k:, 3]).max() / ref[k:, 3].max()) if k < FRAMES else 0.0
        print("  %s: same bubbles for %d of %d frames; there max |target - FP64| / V = %.2e, max |cells - FP64| / V"
              " = %.2e, cells vs target %.2e; after it, total target vs FP64 %.2e, bubbles at the end %d (FP64 %d)"
              % (label, k, FRAMES, dt_, dc, drift, after, r[-1, 2], ref[-1, 2]))
        assert k >= FRAMES // 4, (label, k)
        assert dt_ <= RTOL_TARGET, (label, dt_)
        assert dc <= RTOL_CELLS, (label, dc)
