"""The multigrid preconditioner on the device against the oracle's V-cycle (cicloV,
pressao-cpu.js:241-255): z = M b for random b on the water cells of real states.

This pins every part of the V-cycle at once -- the level masks (tiposGrossos and
prepararVizinhos, pressao-cpu.js:44-81, with labelled air as water), the 16x16 block
red-black smoother and its offsets and colour orders, restriction, prolongation, the
coarsest-level sweeps and the dense operators that replace them -- where a fixed number of PCG
iterations only sees them through a 2e-3 tolerance on q.  The device computes in FP32 (FP16
storage in the mixed mode) against the oracle's binary32-stored FP64 arithmetic; the relative
tolerances below bound that rounding.  Bubble cells are water to the V-cycle; the inputs are
zero there (the device hook fills them from the bubble residual, zero after setup) and the
outputs are compared on the water cells.
"""
import numpy as np
import pytest

from gpu_common import ampulheta, clone_state, dam, hip_sim_like, log, rel_err, substep_dt, tanque_com_bolha

pytestmark = pytest.mark.gpu

AGUA = 1
RTOL = {"fp32": 1e-5, "mixed": 2e-3}   # measured: <= 9.5e-7 (fp32), <= 2.1e-4 (mixed)


@pytest.fixture(scope="module")
def tq():
    import tanque_hip
    tanque_hip.load()
    return tanque_hip


def state(orc, build, frames):
    osim = build(orc)
    for _ in range(frames):
        dt = substep_dt(orc, osim)
        osim.avancar(1 / 60, int(round((1 / 60) / dt)))
    dt = substep_dt(orc, osim)
    for ph in range(8):
        osim.fase(ph, dt)
    return osim


@pytest.mark.parametrize("name,build,frames", [
    ("dam", lambda o: dam(o), 10),                                      # 4 levels, dense V-cycle from level 1
    ("bolha", lambda o: tanque_com_bolha(o), 6),                        # an enclosed bubble
    ("ampulheta", lambda o: ampulheta(o), 12),                          # walls, air pockets
    ("dam256", lambda o: dam(o, nx=256, ny=144, h=1.0 / 144, seed=0), 4),
    # odd row lengths on levels 0 and 1 (161, 81): cell-by-cell tile loads and stores
    ("dam161", lambda o: dam(o, nx=161, ny=90, h=0.8 / 90, seed=0), 4),
    # 1600x800: 8 levels, the last one past the 64x64-tile type pyramid; level 0 over 1M cells
    # runs its down passes as separate kernels
    ("dam1600", lambda o: dam(o, nx=1600, ny=800, h=1.0 / 800, seed=0), 0),
    # 2400x1800: level 1 (1200x900) over 1M cells too, so its up passes are separate kernels
    ("dam2400", lambda o: dam(o, nx=2400, ny=1800, h=1.8 / 1800, seed=0), 0),
])
@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_vcycle_matches_oracle(oracle, tq, name, build, frames, prec):
    osim = state(oracle, build, frames)
    tipo = osim.read(oracle.F_TIPO)
    water = tipo == AGUA
    p = {"fp32": tq.PREC_FP32, "mixed": tq.PREC_MIXED}[prec]
    bad = []
    for dense in (True, False):
        g = hip_sim_like(tq, osim, precision=p, mg_dense=dense)
        clone_state(osim, g)
        rng = np.random.default_rng(7)
        for k in range(3):
            b = np.where(water, rng.standard_normal(water.size), 0.0).astype(np.float32)
            zg = g.vcycle(b)
            zo = osim.vcycle(b.astype(np.float64))
            e = rel_err(zg[water], zo[water])
            log("\n  [%s %s dense=%d] vector %d: z vs oracle %.2e (|z| max %.3e, %d water cells)" % (
                name, prec, dense, k, e, np.abs(zo[water]).max(), int(water.sum())))
            if e > RTOL[prec]:
                bad.append((dense, k, e))
            if np.any(zg[~water] != 0):
                bad.append((dense, k, "nonzero z off the water cells"))
        g.close()
    assert not bad, bad


@pytest.mark.parametrize("name,build,frames", [
    ("dam256", lambda o: dam(o, nx=256, ny=144, h=1.0 / 144, seed=0), 4),
    ("ampulheta", lambda o: ampulheta(o), 12),
    ("dam161", lambda o: dam(o, nx=161, ny=90, h=0.8 / 90, seed=0), 4),
])
@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_split_levels_match_oracle(oracle, tq, name, build, frames, prec):
    """Every tiled level run one pass per kernel (tq_set_launch "split_cells" = 1, a shape the
    autotuner may choose) gives the oracle's V-cycle as the fused kernels do."""
    osim = state(oracle, build, frames)
    water = osim.read(oracle.F_TIPO) == AGUA
    p = {"fp32": tq.PREC_FP32, "mixed": tq.PREC_MIXED}[prec]
    g = hip_sim_like(tq, osim, precision=p)
    clone_state(osim, g)
    g.set_launch("split_cells", 1)
    rng = np.random.default_rng(11)
    for k in range(3):
        b = np.where(water, rng.standard_normal(water.size), 0.0).astype(np.float32)
        zg = g.vcycle(b)
        zo = osim.vcycle(b.astype(np.float64))
        e = rel_err(zg[water], zo[water])
        log("\n  [%s %s split] vector %d: z vs oracle %.2e" % (name, prec, k, e))
        assert e <= RTOL[prec]
    g.close()
