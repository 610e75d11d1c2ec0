"""Run-to-run determinism of the whole substep on the device: two runs of the same hourglass
(a sealed scene with one enclosed bubble, residual-terminated MGPCG) from the same particles
give bit-identical particles, pressures and bubble unknowns.  P2G and the bubble sums use
fixed-point atomics and every solver reduction has a fixed order, so nothing depends on the
order in which wavefronts finish."""
import numpy as np
import pytest

from gpu_common import log

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,graph", [(0, False), (1, True)])
def test_hourglass_run_twice_identical(prec, graph):
    import tanque_hip as tq
    from tanque_hip import scenes
    tq.load()
    work = scenes.scaled_hourglass(224, 126)
    out = []
    for _ in range(2):
        cena = work.cena()
        sim = tq.SimulacaoHIP(cena, tol=1e-6, iteracoes=200, precision=prec, use_graph=graph)
        sim.adicionar(work.particles, False)
        for _ in range(6):
            sim.avancar(work.dt_frame * 4, 4)
        out.append((sim.particulas(), sim.read(12), sim.read(22), sim.read(19), sim.solver_stats()))
        sim.close()
    a, b = out
    log("\n  hourglass 224x126 prec %d graph %d: n %d, iterations %d / %d, bubbles %d" % (
        prec, graph, len(a[0]), a[4]["total_iterations"], b[4]["total_iterations"], a[4]["active_bubbles"]))
    assert a[4]["total_iterations"] == b[4]["total_iterations"]
    for k in range(4):
        assert np.array_equal(a[k], b[k])


@pytest.mark.parametrize("iters", [3, 200])
def test_maze_with_many_bubbles_run_twice_identical(iters):
    """A maze tank traps far more bubbles than the 8 per-block reduction slots; the others are
    summed in 128-bit fixed point, so two runs still agree bit for bit (fixed 3 PCG iterations as
    the reference, and solved to 1e-6)."""
    import tanque_hip as tq
    from tanque_hip import scenes
    from tanque_hip.scenes import passo_frame
    tq.load()
    t = scenes.batched_mazes(n=3, seed=3)[2]
    out = []
    for _ in range(2):
        cena = t.cena()
        sim = tq.SimulacaoHIP(cena, iteracoes=iters, tol=1e-6 if iters > 3 else 0.0)
        sim.adicionar(t.particles, False)
        for _ in range(8):
            sim.avancar(t.dt_frame, passo_frame(t.h, cena.altura, sim.vMax, t.dt_frame))
        out.append((sim.particulas(), sim.read(12), sim.read(22), sim.read(19), sim.solver_stats()))
        sim.close()
    a, b = out
    log("\n  maze %dx%d, %d PCG iterations: n %d, active bubbles %d, iterations %d / %d" % (
        t.nx, t.ny, iters, len(a[0]), a[4]["active_bubbles"], a[4]["total_iterations"], b[4]["total_iterations"]))
    assert a[4]["active_bubbles"] > 8
    assert a[4]["total_iterations"] == b[4]["total_iterations"]
    for k in range(4):
        assert np.array_equal(a[k], b[k])
