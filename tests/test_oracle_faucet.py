"""The seeded faucet (emitirTorneira, principal.js:301-321, with the LCG of testes/apoio.js:4-9 in
place of Math.random) on the oracle: count, disc, velocity clamp, seed advance, acceptance; and
the LCG jump-ahead the device uses to give each particle its own two draws."""
import math

import numpy as np
import pytest

MOLDURA = 2
PPC = 4


def lcg_steps(s, k):
    for _ in range(k):
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
    return s


def lcg_jump(s, k):
    """The device's jump-ahead (tq_device.hpp lcg_jump): powers of the affine map."""
    A, Cc, a, c = 1, 0, 1664525, 1013904223
    while k:
        if k & 1:
            A = (a * A) & 0xFFFFFFFF
            Cc = (a * Cc + c) & 0xFFFFFFFF
        c = (a * c + c) & 0xFFFFFFFF
        a = (a * a) & 0xFFFFFFFF
        k >>= 1
    return (A * s + Cc) & 0xFFFFFFFF


def test_lcg_jump_ahead_equals_sequential_steps():
    rng = np.random.default_rng(3)
    for _ in range(40):
        s = int(rng.integers(0, 2 ** 32))
        k = int(rng.integers(0, 5000))
        assert lcg_jump(s, k) == lcg_steps(s, k)
    assert lcg_jump(12345, 0) == 12345


def expected_count(raio, pvx, pvy, dt, h):
    vx, vy = pvx * 0.5, pvy * 0.5 - 0.7
    v = math.hypot(vx, vy)
    if v > 4:
        vx, vy = vx * 4 / v, vy * 4 / v
    vazao = 2 * raio * max(0.7, math.hypot(vx, vy))
    return min(40000, int(math.floor((vazao * dt) / (h * h) * PPC + 0.5))), (vx, vy)


@pytest.mark.parametrize("pvx,pvy", [(0.0, -1.6), (0.3, 0.2), (30.0, -40.0)])
def test_faucet_on_the_oracle(oracle, pvx, pvy):
    h = 0.01
    cena = oracle.Cena(80, 60, h)
    sim = oracle.Simulacao(cena)
    raio, dt = 0.06, 1 / 60
    x, y = 0.4, 0.45
    sem = oracle.Semente(17)
    s0 = sem.state.value
    n, acc = sim.emitir_torneira(x, y, pvx, pvy, raio, (0.1, 0.2, 0.3), dt, sem)
    n_exp, (vx, vy) = expected_count(raio, pvx, pvy, dt, h)
    assert n == n_exp > 0
    assert sem.state.value == lcg_steps(s0, 2 * n)
    P = sim.particulas().astype(np.float64)
    assert len(P) == acc <= n
    assert acc > 0.9 * n  # an empty tank, away from the walls: every cell starts empty
    d = np.hypot(P[:, 0] - x, P[:, 1] - y)
    assert d.max() <= raio * (1 + 1e-6)
    assert np.all(P[:, 2] == np.float32(vx)) and np.all(P[:, 3] == np.float32(vy))
    assert math.hypot(vx, vy) <= 4 + 1e-12
    assert np.all(P[:, 9] == np.float32(0.1)) and np.all(P[:, 11] == np.float32(0.3))
    # the first particle from the first two draws, as the reference builds it
    u1 = lcg_steps(s0, 1) / 2 ** 32
    u2 = lcg_steps(s0, 2) / 2 ** 32
    a, r = u1 * math.pi * 2, math.sqrt(u2) * raio
    assert P[0, 0] == np.float32(x + math.cos(a) * r) and P[0, 1] == np.float32(y + math.sin(a) * r)
