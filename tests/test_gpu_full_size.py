"""The Ultra configuration at its full size (sealed hourglass 896x504, 745,943 particles, MGPCG to a
1e-6 relative residual), through properties that do not depend on size: every solve of every
frame reaches the tolerance within the iteration cap, the sealed tank keeps every particle and
every particle stays inside the tank, the lower chamber's air is one tracked bubble whose target
volume (volumeAlvo, ar.js) is conserved while it persists, and two FP32 runs are bit-identical.
The mixed-precision preconditioner is held to the same properties, and the Alto configuration
(cup and faucet, 480x270) to the matching ones."""
import numpy as np
import pytest

from gpu_common import log

pytestmark = pytest.mark.gpu

FRAMES = 24   # 1/20 s at 1/480 s
TOL = 1e-6
MAXIT = 200


def run(tq, prec):
    from tanque_hip import scenes
    from tanque_hip.scenes import passo_frame
    w = scenes.ultra_hourglass()
    cena = w.cena()
    sim = tq.SimulacaoHIP(cena, precision=prec, tol=TOL, iteracoes=MAXIT)
    n0 = sim.adicionar(w.particles, False)
    W, H = w.nx * w.h, w.ny * w.h
    worst, alvo = 0.0, []
    for _ in range(FRAMES):
        sim.avancar(w.dt_frame, passo_frame(w.h, H, sim.vMax, w.dt_frame))
        st = sim.solver_stats()
        assert st["iterations"] < MAXIT, st
        worst = max(worst, st["norm"] / st["norm_b"])
        va = np.asarray(sim.read(tq.F_VALVO), np.float64)
        alvo.append((int((va[2:] > 0).sum()), float(va[2:].sum())))
    P = sim.particulas()
    out = (P, sim.read(tq.F_Q), sim.read(tq.F_YBOLHA))
    n = sim.n
    sim.close()
    return n0, n, P, W, H, worst, alvo, out


@pytest.fixture(scope="module")
def tq():
    import tanque_hip
    tanque_hip.load()
    return tanque_hip


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_ultra_full_size_properties(tq, prec):
    p = tq.PREC_FP32 if prec == "fp32" else tq.PREC_MIXED
    n0, n, P, W, H, worst, alvo, _ = run(tq, p)
    log("\n  ultra %s: %d particles kept of %d, worst last-solve residual %.2e, bubbles %s, target %.1f -> %.1f"
        % (prec, n, n0, worst, sorted({b for b, _ in alvo}), alvo[0][1], alvo[-1][1]))
    assert n0 == 745943
    assert n == n0 and len(P) == n0                     # sealed: nothing leaves
    assert worst <= TOL
    assert np.all(np.isfinite(P[:, :4]))
    assert np.all((P[:, 0] > 0) & (P[:, 0] < W) & (P[:, 1] > 0) & (P[:, 1] < H))
    assert all(b == 1 for b, _ in alvo)                 # the lower chamber's one bubble
    v = np.array([s for _, s in alvo])
    assert np.max(np.abs(v - v[0])) <= 1e-2 * v[0], v   # its target volume persists


def test_ultra_full_size_run_twice_identical(tq):
    a = run(tq, tq.PREC_FP32)[-1]
    b = run(tq, tq.PREC_FP32)[-1]
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def run_alto(tq, frames):
    from tanque_hip import scenes
    from tanque_hip.scenes import passo_frame
    w = scenes.alto_cup()
    f = w.extra["faucet"]
    cena = w.cena()
    sim = tq.SimulacaoHIP(cena, tol=TOL, iteracoes=MAXIT)
    sem = tq.Semente(f["seed"])
    emitted = kept = 0
    worst = 0.0
    for k in range(frames):
        if k < f["frames"]:
            n, acc = sim.emitirTorneira(f["x"], f["y"], f["pvx"], f["pvy"], f["raio"], f["absorcao"], w.dt_frame, sem)
            assert 0 <= acc <= n
            emitted += n
            kept += acc
        sim.avancar(w.dt_frame, passo_frame(w.h, cena.altura, sim.vMax, w.dt_frame))
        st = sim.solver_stats()
        if st["norm_b"] > 0:
            assert st["iterations"] < MAXIT, st
            worst = max(worst, st["norm"] / st["norm_b"])
    P = sim.particulas()
    n = sim.n
    sim.close()
    return emitted, kept, n, P, w.nx * w.h, w.ny * w.h, worst


def test_alto_full_size_properties(tq):
    """The Alto configuration at full size (480x270 cup, faucet 12 cells wide at 1.5 m/s), its
    first 120 frames: every solve reaches 1e-6, the particles are exactly those the faucet's adds
    kept (none is removed), all inside the tank, and a rerun is bit-identical."""
    a = run_alto(tq, 120)
    emitted, kept, n, P, W, H, worst = a
    log("\n  alto: %d emitted, %d kept, %d at the end, worst last-solve residual %.2e" % (emitted, kept, n, worst))
    assert emitted > 0 and 0 < kept <= emitted
    assert n == kept
    assert worst <= TOL
    assert np.all(np.isfinite(P[:, :4]))
    assert np.all((P[:, 0] > 0) & (P[:, 0] < W) & (P[:, 1] > 0) & (P[:, 1] < H))
    b = run_alto(tq, 120)
    assert a[:3] == b[:3]
    assert np.array_equal(a[3], b[3])
