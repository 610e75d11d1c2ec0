"""The oracle driven from Node (oracle/bench_oracle.js over oracle/oracle_node.node, the CPU
baseline's driver) runs the same frames as the oracle driven from Python: same accepted particle
count and the same substep schedule, frame by frame (each frame's count comes from the top speed
the previous frames left, so a different state would change it)."""
import json
import math
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


@pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")
def test_node_driver_matches_python_driver(tmp_path):
    from oracle import oracle as orc
    orc.build()
    if not os.path.exists(orc.NODE_ADDON):
        pytest.skip("node headers absent: the oracle's Node addon is not built")
    nx, ny, h = 96, 54, 0.8 / 54
    strokes = [(0.1, 0.1, 0.6, 0.35, 0.02, False), (0.9, 0.5, 1.2, 0.5, 0.03, False), (0.9, 0.5, 1.0, 0.5, 0.01, True)]
    cena = orc.Cena(nx, ny, h)
    for s in strokes:
        cena.tracar(*s)
    P = orc.semear_retangulo(2 * h, 2 * h, 0.5, 0.4, h, (0, 0, 0), orc.Semente(7))
    pf = tmp_path / "p.f32"
    np.ascontiguousarray(P, np.float32).tofile(pf)
    frames, dt = 6, 1 / 120
    wf = tmp_path / "w.json"
    wf.write_text(json.dumps({"nx": nx, "ny": ny, "h": h, "strokes": [list(s) for s in strokes],
                              "particles_file": str(pf), "dt_frame": dt, "frames": frames, "budget_s": 600,
                              "tol": 1e-6, "maxit": 200}))
    r = subprocess.run(["node", os.path.join(ROOT, "oracle", "bench_oracle.js"), str(wf)], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    sim = orc.Simulacao(cena, mode=0, tol=1e-6, iteracoes=200)
    n = sim.adicionar(P, False)
    subs = 0
    for _ in range(frames):
        sub = int(math.ceil(dt / orc.passo_maximo(h, cena.altura, sim.vmax, dt)))
        sim.avancar(dt, sub)
        subs += sub
    assert out["particles"] == n > 0
    assert out["frames"] == frames
    assert out["substeps"] == subs
