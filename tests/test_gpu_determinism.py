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
