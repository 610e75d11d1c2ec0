"""tq_water_fraction_noalloc (the entry point that allocates nothing) against tq_water_fraction,
through the C ABI: two calls in a row on changing states, the second with a smaller box, each
within 1e-12 of the allocating entry point (both sum FP64 density splats added by atomics, so
their last bits can differ from call to call), a repeated call on one state within the same
1e-12 (a scratch left uncleared by the first call would add the whole fraction again), and boxes
beyond the scratch refused with TQ_E_CAPACITY."""
import ctypes as C

import pytest

from gpu_common import dam, hip_sim_like, log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tq():
    import tanque_hip
    tanque_hip.load()
    return tanque_hip


def noalloc(tq, sim, box):
    f = C.c_double(-1.0)
    rc = tq.load().tq_water_fraction_noalloc(sim._h, *[int(v) for v in box], C.byref(f))
    return rc, f.value


def test_noalloc_matches_water_fraction(oracle, tq):
    osim = dam(oracle, nx=160, ny=90, h=0.8 / 90, seed=0)
    g = hip_sim_like(tq, osim)
    g.write_particles(osim.particulas())
    boxes = [(2, 2, 158, 88), (10, 2, 60, 40)]  # the second call: a smaller box, inside the water
    for k, box in enumerate(boxes):
        g.avancar(1 / 60, 4)
        ref = g.water_fraction(box)
        rc, val = noalloc(tq, g, box)
        assert rc == 0, tq.load().tq_last_error()
        rc2, again = noalloc(tq, g, box)
        assert rc2 == 0
        log("\n  call %d box %s: noalloc %.15f, tq_water_fraction %.15f, repeat %.15f" % (k, box, val, ref, again))
        assert abs(val - ref) <= 1e-12 * max(1.0, abs(ref))
        assert abs(again - val) <= 1e-12 * max(1.0, abs(val))
        assert 0.0 < ref <= 1.0


def test_noalloc_refuses_large_box(tq, oracle):
    osim = dam(oracle, nx=1100, ny=960, h=0.8 / 960, seed=0)
    g = hip_sim_like(tq, osim)
    rc, _ = noalloc(tq, g, (0, 0, 1100, 960))  # 1,056,000 cells > 1,048,576
    assert rc == 3  # TQ_E_CAPACITY
    assert b"scratch" in tq.load().tq_last_error()
    rc, _ = noalloc(tq, g, (0, 0, 1000, 960))  # 960,000 cells fit
    assert rc == 0
    rc, _ = noalloc(tq, g, (5, 5, 5, 10))  # empty box
    assert rc == 1  # TQ_E_ARG
