"""Batched tanks (tq_batch_*): stepping tanks together gives, tank by tank, exactly the states of
stepping each alone, bit for bit, in both launch models -- one table launch per kernel over all
tanks (each block runs as its tank's own logical block) and every tank on its own stream fed by
host threads -- with air tracking on: a maze traps 100-200 bubbles, and the bubbles beyond the 8
per-block reduction slots are summed in order-independent fixed point.  The substep counts come
from every tank's vMax after the frame (tq_batch_vmax for the batch, tq_vmax alone), so they
match only if the gathered vMax does."""
import numpy as np
import pytest

from gpu_common import log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tq():
    import tanque_hip
    tanque_hip.load()
    return tanque_hip


def frame_subs(ts, cenas, vmaxes):
    from tanque_hip.scenes import passo_frame
    return [passo_frame(t.h, c.altura, v, t.dt_frame) for t, c, v in zip(ts, cenas, vmaxes)]


@pytest.mark.parametrize("mode", ["table", "table-graph", "threads"])
@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_batch_matches_tanks_stepped_alone(tq, prec, mode):
    """Frame by frame: table-graph captures the first frame's table launches as a HIP graph and
    replays it for the frames with the same substep counts, with host reads between the replays
    (table launches every table kernel itself)."""
    from tanque_hip import scenes
    p = tq.PREC_FP32 if prec == "fp32" else tq.PREC_MIXED
    ts = scenes.batched_mazes(n=5, seed=3)
    opts = {"iteracoes": 3, "precision": p}
    cenas = [t.cena() for t in ts]
    b = tq.Batch(cenas, [opts] * len(ts), threads=3, mode=mode.split("-")[0])
    if mode == "table-graph":
        b.use_graphs(True)
    alone = [tq.SimulacaoHIP(t.cena(), iteracoes=3, precision=p) for t in ts]
    for t, sb, sa in zip(ts, b.tanks, alone):
        sb.adicionar(t.particles, False)
        sa.adicionar(t.particles, False)
    for f in range(6):
        subs = frame_subs(ts, cenas, b.vmax())
        assert subs == frame_subs(ts, cenas, [s.vMax for s in alone])
        b.step(ts[0].dt_frame, subs)
        for k, (sa, sb, ns) in enumerate(zip(alone, b.tanks, subs)):
            sa.avancar(ts[0].dt_frame, ns)
            assert np.array_equal(sb.particulas().view(np.uint32), sa.particulas().view(np.uint32)), (f, k)
    if mode == "table-graph":
        st = b.stats()
        log("\n  frame graphs: %d captured, %d replayed" % (st["graph_builds"], st["graph_replays"]))
        assert st["graph_replays"] >= 1
    for k, (t, sb, sa) in enumerate(zip(ts, b.tanks, alone)):
        pb, pa = sb.particulas(), sa.particulas()
        log("\n  %s %s tank %d (%dx%d): %d particles batched, %d alone" % (mode, prec, k, t.nx, t.ny, len(pb),
                                                                           len(pa)))
        assert len(pb) == len(pa) > 0
        assert np.array_equal(pb.view(np.uint32), pa.view(np.uint32))
        assert np.array_equal(sb.read(12), sa.read(12))
    for s in alone:
        s.close()
    b.close()


@pytest.mark.parametrize("mode", ["table", "threads"])
def test_batch_rejects_bad_arguments(tq, mode):
    from tanque_hip import scenes
    ts = scenes.batched_mazes(n=2, seed=4)
    b = tq.Batch([t.cena() for t in ts], [{}] * 2, threads=2, mode=mode)
    with pytest.raises(tq.TqError):
        b.step(-1.0, [1, 1])
    with pytest.raises(tq.TqError):
        b.step(1 / 240, [1, -1])
    b.close()


def test_table_batch_needs_fixed_iterations(tq):
    """A table frame cannot poll a residual-terminated solve: a tank with tol > 0 is refused."""
    from tanque_hip import scenes
    ts = scenes.batched_mazes(n=2, seed=4)
    b = tq.Batch([t.cena() for t in ts], [{}, {"tol": 1e-6, "iteracoes": 50}], threads=2, mode="table")
    with pytest.raises(tq.TqError, match="fixed PCG iteration count"):
        b.step(1 / 240, [1, 1])
    b.close()


@pytest.mark.parametrize("mode", ["table", "threads"])
def test_full_batched_configuration_matches_tanks_stepped_alone(tq, mode):
    """The configured batch: 64 maze tanks of uniform random sizes (64-384 x 48-256 cells, 2.39 M
    cells and 2.07 M particles in all), 3 PCG iterations as the reference; three frames batched
    equal the same frames stepped tank after tank, bit for bit."""
    from tanque_hip import scenes
    ts = scenes.batched_mazes()
    assert len(ts) == 64
    cenas = [t.cena() for t in ts]
    b = tq.Batch(cenas, [{"iteracoes": 3}] * len(ts), threads=8, mode=mode)
    alone = [tq.SimulacaoHIP(t.cena(), iteracoes=3) for t in ts]
    for t, sb, sa in zip(ts, b.tanks, alone):
        sb.adicionar(t.particles, False)
        sa.adicionar(t.particles, False)
    for f in range(3):
        subs = frame_subs(ts, cenas, b.vmax())
        b.step(ts[0].dt_frame, subs)
        for sa, k in zip(alone, subs):
            sa.avancar(ts[0].dt_frame, k)
    cells = parts = 0
    for k, (t, sb, sa) in enumerate(zip(ts, b.tanks, alone)):
        pb, pa = sb.particulas(), sa.particulas()
        assert len(pb) == len(pa) > 0, k
        assert np.array_equal(pb.view(np.uint32), pa.view(np.uint32)), k
        assert np.array_equal(sb.read(12), sa.read(12)), k
        cells += t.nx * t.ny
        parts += len(pb)
    log("\n  %s: 64 tanks, %d cells, %d particles: batched == alone after 3 frames" % (mode, cells, parts))
    for s in alone:
        s.close()
    b.close()
