"""The Node-API addon (the JavaScript host's binding of the C ABI) and its JS wrappers
(tanque-de-agua_amd/js/tanque_hip.js): host helpers on the CPU, and on the GPU a dam break
driven through MotorHIP in node gives particles bit-identical to the same run through the
ctypes binding (same library, deterministic kernels)."""
import base64
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "tanque-de-agua_amd", "lib", "tanque_node.node")
DRIVER = os.path.join(ROOT, "tests", "js", "run_dam.js")

needs_node = pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(ADDON),
                                reason="node or the built addon is absent")


def run_node(*args, timeout=300):
    r = subprocess.run(["node", DRIVER] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    d["P"] = np.frombuffer(base64.b64decode(d.pop("particulas")), np.float32).reshape(-1, 12)
    return d


@needs_node
def test_addon_host_helpers_match_the_library():
    import tanque_hip as tq
    tq.load()
    h = 0.01
    d = run_node(80, 60, h, 0.1, "host-only")
    sem = tq.Semente(9)
    P = tq.semear_retangulo(2 * h, 2 * h, 0.3, 0.42, h, (0, 0, 0), sem).reshape(-1, 12)
    assert np.array_equal(d["P"], P)
    assert d["semente"] == sem.state.value
    assert d["passo"] == tq.passo_maximo(h, 60 * h, None, 1 / 60)


@needs_node
@pytest.mark.gpu
def test_motor_hip_in_node_matches_ctypes_run():
    import tanque_hip as tq
    tq.load()
    nx, ny, h, seconds = 80, 60, 0.01, 0.5
    d = run_node(nx, ny, h, seconds)
    cena = tq.Cena(nx, ny, h)
    sim = tq.SimulacaoHIP(cena)
    sim.adicionar(tq.semear_retangulo(2 * h, 2 * h, 0.3, 0.42, h, (0, 0, 0), tq.Semente(9)), False)
    quadro = 1 / 60
    for _ in range(int(math.floor(seconds / quadro + 0.5))):
        sub = math.ceil(quadro / tq.passo_maximo(h, cena.altura, sim.vMax, quadro))
        sim.avancar(quadro, sub)
    P = sim.particulas()
    print("\n  node MotorHIP: n %d, vMax %.4f, %.3f ms/substep; ctypes n %d" % (
        d["n"], d["vMax"], d["custoPorSubpasso"], len(P)))
    assert d["n"] == len(P)
    assert np.array_equal(d["P"], P)
    assert d["vMax"] == sim.vMax


@needs_node
@pytest.mark.gpu
def test_node_render_water_matches_ctypes():
    """renderizarAgua() through the addon returns the images render_water() returns through
    ctypes for the same run (sizes 3 nx x 3 ny RGBA; bytes equal up to a one-unit rounding edge
    where the per-cell particle order, which the sort's atomics set, changes a binary64 sum)."""
    import tanque_hip as tq
    tq.load()
    nx, ny, h, seconds = 80, 60, 0.01, 0.25
    r = subprocess.run(["node", DRIVER] + [str(v) for v in (nx, ny, h, seconds, "render")], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    assert (d["largura"], d["altura"]) == (3 * nx, 3 * ny)
    tint = np.frombuffer(base64.b64decode(d["tinta"]), np.uint8).reshape(3 * ny, 3 * nx, 4)
    foam = np.frombuffer(base64.b64decode(d["espuma"]), np.uint8).reshape(3 * ny, 3 * nx, 4)
    cena = tq.Cena(nx, ny, h)
    sim = tq.SimulacaoHIP(cena)
    sim.adicionar(tq.semear_retangulo(2 * h, 2 * h, 0.3, 0.42, h, (0, 0, 0), tq.Semente(9)), False)
    quadro = 1 / 60
    for _ in range(int(math.floor(seconds / quadro + 0.5))):
        sim.avancar(quadro, math.ceil(quadro / tq.passo_maximo(h, cena.altura, sim.vMax, quadro)))
    t2, f2 = sim.render_water()
    for a, b in ((tint, t2), (foam, f2)):
        diff = np.abs(a.astype(int) - b.astype(int))
        assert diff.max() <= 1 and (diff > 0).sum() <= 1e-3 * diff.size
    assert (tint[..., 0] < 250).any()


@needs_node
@pytest.mark.gpu
def test_node_backward_matches_ctypes_backward():
    """The funnel's gradient through the addon (gravar, fracaoAgua, retropropagar) equals the
    ctypes run's: same library and decisions.  The reverse pass's FP64 atomic sums are taken in
    no fixed order and the adjoint solve's right-hand side is stored in FP32, so the two differ
    by a few FP32 roundings (1e-5 relative here)."""
    import tanque_hip as tq
    from tanque_hip import scenes
    from tanque_hip.adjunto import run_funnel
    tq.load()
    w = scenes.funnel(nx=96, ny=54, frames=40)
    box = (30, 2, 94, 52)
    w.extra["box"] = box
    f = w.extra["faucet"]
    cfg = {"nx": w.nx, "ny": w.ny, "h": w.h, "strokes": [list(s) for s in w.strokes], "frames": w.frames,
           "dt": w.dt_frame, "box": list(box), "every": 10,
           "faucet": {"x": f["x"], "y": f["y"], "pvx": f["pvx"], "pvy": f["pvy"], "raio": f["raio"], "seed": f["seed"]}}
    script = os.path.join(ROOT, "tests", "js", "run_funnel_adjoint.js")
    r = subprocess.run(["node", script, json.dumps(cfg)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    out = run_funnel(w, checkpoint_every=10)
    print("\n  node: frac %.9g jet %s; ctypes: frac %.9g jet %s" % (d["fracao"], d["jato"], out["frac"], out["g_vjet"]))
    assert d["fracao"] == out["frac"]
    assert d["chamadas"] == w.frames
    np.testing.assert_allclose(d["jato"], out["g_vjet"], rtol=1e-5, atol=1e-5 * np.abs(out["g_vjet"]).max())
    gs = np.frombuffer(base64.b64decode(d["sdf"]), np.float32)
    assert np.linalg.norm(gs - out["g_sdf"]) <= 1e-5 * np.linalg.norm(out["g_sdf"]) + 1e-30


@needs_node
@pytest.mark.gpu
def test_motor_hip_asynchronous_frames():
    """MotorHIP as MotorGPU (motor.js:54-74): 30 frames queued without waiting, then contagem()
    resolves to the live count and vMax/custoPorSubpasso are those of the last frame, equal to
    the blocking reads of the same state; lists beyond maxParticulas are cut to the room left
    (motor.js:100-104), so the tank fills exactly to the cap and stays there."""
    import tanque_hip as tq
    tq.load()
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "check_motor.js")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    print("\n  30 frames queued in %.1f ms host time, count back after %.1f ms (ready at once: %s); n %d, vMax %.4f, "
          "%.4f ms/substep; cap %d -> %d, %d" % (d["msFila"], d["msAteContagem"], d["prontoLogo"], d["n"], d["vMax"],
                                                d["custo"], d["teto"], d["nNoTeto"], d["nDepois"]))
    assert d["vMaxAntes"] is None
    assert d["n"] == d["nBloqueante"] > 0
    assert d["vMax"] == d["vMaxBloqueante"] and d["vMax"] > 0
    assert d["custo"] is not None and 0 < d["custo"] < 1000
    assert d["teto"] == math.ceil(240 * 135 * 4.6)
    assert d["nNoTeto"] == d["teto"] == d["nDepois"] == d["nParticulas"]
