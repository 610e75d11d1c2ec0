"""The render hand-off (tq_render_water) against a restatement of the reference's 2D renderer's
water layer (js/render-2d.js:110-162) in numpy: every particle a (1 - r^2)^2 splat of radius
2.1 cells at 3 pixels per cell into coverage, foam and dye, added in the reference's order
(particle, then pixel row, then pixel column) into Float32Array buffers as the reference does
(each add in binary64, the sum rounded to binary32), then the Beer-Lambert tint and the foam alpha
stored as a Uint8ClampedArray stores them (round half to even, clamp).  The device gathers per
pixel in another order, sums in binary64 and uses the device exp, so a byte may differ by one
where a value sits on a rounding edge; the test bounds how many.  Rendering between frames
leaves the simulation bit-identical."""
import math

import numpy as np
import pytest

from gpu_common import dam, log

pytestmark = pytest.mark.gpu

PX, RAIO, PROF, PPC = 3, 2.1, 0.05, 4
ABS = (9.0, 2.6, 1.6)


def suave(a, b, x):
    t = np.minimum(1.0, np.maximum(0.0, (x - a) / (b - a)))
    return t * t * (3 - 2 * t)


def render_reference(P, nx, ny, h):
    w, hh = nx * PX, ny * PX
    escala = PX / h
    R = RAIO * PX
    R2 = R * R
    cx = P[:, 0].astype(np.float64) * escala
    cy = hh - P[:, 1].astype(np.float64) * escala
    x0 = np.maximum(0, np.ceil(cx - R - 0.5)).astype(np.int64)
    x1 = np.minimum(w - 1, np.floor(cx + R - 0.5)).astype(np.int64)
    y0 = np.maximum(0, np.ceil(cy - R - 0.5)).astype(np.int64)
    y1 = np.minimum(hh - 1, np.floor(cy + R - 0.5)).astype(np.int64)
    span = int(math.ceil(2 * R)) + 2
    oy, ox = np.meshgrid(np.arange(span), np.arange(span), indexing="ij")   # row-major: y, then x
    X = x0[:, None, None] + ox[None]
    Y = y0[:, None, None] + oy[None]
    inside = (X <= x1[:, None, None]) & (Y <= y1[:, None, None])
    dx = X + 0.5 - cx[:, None, None]
    dy = Y + 0.5 - cy[:, None, None]
    r2 = (dx * dx + dy * dy) / R2
    ok = inside & (r2 < 1)
    peso = np.where(ok, (1 - r2) * (1 - r2), 0.0)
    idx = np.where(ok, Y * w + X, 0).ravel()
    sel = ok.ravel()
    idx, pw = idx[sel], peso.ravel()[sel]
    rep = np.repeat(np.arange(len(P)), span * span)[sel]
    # Float32Array sums in the reference's order: per pixel, its addends in particle order (the
    # stable sort keeps it), the j-th addend of every pixel added in step j
    order = np.argsort(idx, kind="stable")
    idx, pw, rep = idx[order], pw[order], rep[order]
    first = np.r_[0, np.flatnonzero(np.diff(idx)) + 1]
    nth = np.arange(len(idx)) - np.repeat(first, np.diff(np.r_[first, len(idx)]))
    adds = [pw, pw * P[rep, 8].astype(np.float64)] + [pw * P[rep, 9 + c].astype(np.float64) for c in range(3)]
    acc = np.zeros((5, w * hh), np.float32)
    for j in range(int(nth.max()) + 1 if len(nth) else 0):
        m = nth == j
        for q in range(5):
            acc[q, idx[m]] = (acc[q, idx[m]].astype(np.float64) + adds[q][m]).astype(np.float32)
    campo = acc[0].astype(np.float64)
    espuma = acc[1].astype(np.float64)
    cor = acc[2:].astype(np.float64)
    phi0 = (PPC * math.pi * RAIO * RAIO) / 3
    phi = campo / phi0
    cobre = suave(0.26, 0.5, phi)
    esp = suave(0.15, 0.95, phi) * PROF
    inv = 1 / np.maximum(campo, 1e-4)
    borda = 1 - cobre * (1 - cobre) * 4 * 0.45
    u8 = lambda v: np.clip(np.rint(v), 0, 255).astype(np.uint8)
    tint = np.empty((w * hh, 4), np.uint8)
    for c in range(3):
        t = np.exp(-(ABS[c] + cor[c] * inv) * esp) * borda
        tint[:, c] = u8((1 - cobre + cobre * t) * 255)
    tint[:, 3] = 255
    e = suave(0.05, 0.75, np.minimum(1.0, espuma * inv)) * 0.85 * cobre
    foam_a = u8(e * 255)
    covered = cobre >= 0.002
    tint[~covered] = 255
    foam_a[~covered] = 0
    return tint.reshape(hh, w, 4), foam_a.reshape(hh, w), covered.reshape(hh, w)


@pytest.fixture(scope="module")
def tq():
    import tanque_hip
    tanque_hip.load()
    return tanque_hip


def scene(tq, oracle):
    from gpu_common import hip_sim_like
    osim = dam(oracle)
    sim = hip_sim_like(tq, osim)
    sim.write_particles(osim.particulas())
    return sim


def test_render_water_matches_reference_layer(tq, oracle):
    sim = scene(tq, oracle)
    for f in range(12):
        if f == 4:  # dye part of the water, as the dye tool does
            sim.aplicarFerramenta({"tipo": "corante", "x": 0.15, "y": 0.2, "raio": 0.1, "absorcao": (0.0, 14.0, 20.0)})
        sim.avancar(1 / 60, 4)
    P = sim.particulas()
    tint, foam = sim.render_water()
    rt, rfa, covered = render_reference(P, sim.nx, sim.ny, sim.h)
    assert tint.shape == rt.shape and foam.shape[:2] == rfa.shape
    dt = np.abs(tint.astype(int) - rt.astype(int))
    df = np.abs(foam[..., 3].astype(int) - rfa.astype(int))
    log("\n  render %dx%d px, %d particles: covered %.1f %%, foam max %d, tint bytes off by one %d, foam %d" % (
        tint.shape[1], tint.shape[0], len(P), 100 * covered.mean(), int(rfa.max()), int((dt > 0).sum()),
        int((df > 0).sum())))
    assert covered.mean() > 0.05 and (rt[covered][:, 0] < 250).any()
    assert dt.max() <= 1 and df.max() <= 1
    assert (dt > 0).sum() <= 1e-3 * dt.size and (df > 0).sum() <= 1e-3 * df.size
    assert np.all(foam[covered][:, :3] == np.array([248, 251, 252], np.uint8))
    assert np.all(foam[~covered] == 0)
    sim.close()


def test_render_leaves_the_simulation_unchanged(tq, oracle):
    a, b = scene(tq, oracle), scene(tq, oracle)
    for _ in range(6):
        a.avancar(1 / 60, 4)
        b.avancar(1 / 60, 4)
        b.render_water()
    assert np.array_equal(a.particulas().view(np.uint32), b.particulas().view(np.uint32))
    assert np.array_equal(a.read(12), b.read(12))
    a.close()
    b.close()
