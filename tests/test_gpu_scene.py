"""Walls built on the device (tq_set_strokes) against the host builder (tq_scene_stroke, itself
bit-identical to the oracle's Cena), and the stroke chain rule (tq_strokes_vjp) against the
host's vertical-shift chain rule and against differences of a binary64 SDF builder."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tanque-de-agua_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

pytestmark = pytest.mark.gpu


def scenes_to_check():
    from tanque_hip import scenes
    rng = np.random.default_rng(3)
    out = []
    w = scenes.funnel(heights=rng.uniform(-0.01, 0.01, 2 * scenes.FUNNEL_SEGMENTS))
    out.append(("funnel", w.nx, w.ny, w.h, w.strokes))
    a = scenes.alto_cup()
    out.append(("alto", a.nx, a.ny, a.h, a.strokes))
    for t in scenes.batched_mazes(n=3):
        out.append(("maze", t.nx, t.ny, t.h, t.strokes))
    # random strokes with erasing, some outside the tank
    nx, ny, h = 120, 70, 0.01
    st = []
    for k in range(40):
        x0, y0, x1, y1 = rng.uniform(-0.1, 1.3), rng.uniform(-0.1, 0.8), rng.uniform(-0.1, 1.3), rng.uniform(-0.1, 0.8)
        st.append((x0, y0, x1, y1, rng.uniform(0.002, 0.06), bool(k % 5 == 4)))
    st.append((0.3, 0.3, 0.3, 0.3, 0.05, False))  # a point stroke
    st.append((5.0, 5.0, 6.0, 6.0, 0.01, False))  # touches no lattice point
    out.append(("random", nx, ny, h, st))
    return out


def test_device_walls_equal_host_walls():
    import tanque_hip as tq
    for name, nx, ny, h, strokes in scenes_to_check():
        cena = tq.Cena(nx, ny, h)
        for s in strokes:
            cena.tracar(*s)
        sim = tq.SimulacaoHIP(tq.Cena(nx, ny, h))
        sim.definir_tracos(strokes)
        got = sim.ler_sdf()
        assert np.array_equal(got, cena.sdf), (name, int(np.sum(got != cena.sdf)))
        ref = tq.SimulacaoHIP(cena)
        assert np.array_equal(sim.solidoCelula, ref.solidoCelula), name
        sim.close()
        ref.close()


def test_stroke_chain_rule():
    import tanque_hip as tq
    from tanque_hip import scenes
    from test_oracle_adjoint import sdf_f64
    nx, ny = 96, 54
    h = scenes.TANK_H / ny
    rng = np.random.default_rng(5)
    theta = rng.uniform(-0.01, 0.01, 2 * scenes.FUNNEL_SEGMENTS)
    strokes = scenes.funnel_strokes(theta, nx, ny)
    g = rng.standard_normal(4 * nx * ny).astype(np.float32)
    sim = tq.SimulacaoHIP(tq.Cena(nx, ny, h))
    d = sim.tracos_vjp(strokes, g)
    host = scenes.stroke_shift_vjp(nx, ny, h, strokes, g.astype(np.float64), list(range(len(strokes))))
    np.testing.assert_allclose(d[:, 1] + d[:, 3], host, rtol=1e-9, atol=1e-9 * np.abs(host).max())
    # each of the five parameters against differences of the binary64 builder
    eps = 1e-7
    for k in (0, 7, 16, 30):
        for q in range(5):
            sp = [list(s) for s in strokes]
            sm = [list(s) for s in strokes]
            sp[k][q] += eps
            sm[k][q] -= eps
            fd = (sdf_f64(nx, ny, h, sp) @ g - sdf_f64(nx, ny, h, sm) @ g) / (2 * eps)
            assert abs(fd - d[k, q]) <= 1e-4 * max(abs(fd), abs(d[k, q])) + 1e-6 * np.abs(d).max(), (k, q, fd, d[k, q])
    sim.close()
