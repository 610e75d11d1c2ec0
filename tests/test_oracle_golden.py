"""The JS-exact oracle against golden vectors of the reference's own CPU engine
(tests/golden/reference_runs.json, written by tests/golden/gen_reference_runs.py from the
reference's JavaScript run under Node): the same seeded scenes of the reference's models, frame
by frame and phase by phase, must give bit-identical state after every phase -- every grid
array, label, bubble array, the tracker's volumes and live flags, and the particles -- and
identical particles at the checkpoints.

This pins, bit for bit over whole runs, what no closed form pins: the multigrid smoother and
V-cycle, the 3-iteration warm-started PCG, #superficie, the tracker's label and volume history
and the G2P collision trajectories; one run adds the tools (dye, hand, sponge) and water added
with the full-cell check between frames (adicionar's accepted counts included).  The foam
column is compared on its own: foam uses Math.exp (V8's fdlibm port) where the oracle uses the
C library's exp.
"""
import base64
import hashlib
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
FIXTURE = os.path.join(ROOT, "tests", "golden", "reference_runs.json")

GRID_FIELDS = list(range(25))   # F_U .. F_VIVO, the order of the reference run's record
PHASES = ["P2G", "GRAV", "TIPOS", "ROTULOS", "SUPERFICIE", "PAR", "CLASSIFICAR", "BOLHAS", "PRESSAO", "GRADIENTE",
          "EXTRAPOLAR", "G2P"]


def sha(bufs):
    h = hashlib.sha1()
    for b in bufs:
        h.update(b)
    return h.hexdigest()[:16]


def record(orc, sim):
    bufs = [np.asarray(sim.read(f), np.float64).tobytes() for f in GRID_FIELDS]
    P = sim.particulas().astype(np.float64)
    bufs.append(np.delete(P, 8, axis=1).tobytes())
    return {"s": sha(bufs), "f": sha([P[:, 8].copy().tobytes()]), "n": len(P)}


def load():
    with open(FIXTURE) as f:
        return json.load(f)


SCENES = [k for k in load() if not k.startswith("_")] if os.path.exists(FIXTURE) else []


@pytest.mark.parametrize("nome", SCENES)
def test_oracle_reproduces_the_reference_run(nome):
    from oracle import oracle as orc
    run = load()[nome]
    cfg = run["config"]
    nx, ny, h = cfg["nx"], cfg["ny"], cfg["h"]
    cena = orc.Cena(nx, ny, h)
    cena.set_sdf(np.frombuffer(base64.b64decode(run["sdf"]), np.float32))
    op = cfg["opcoes"]
    sim = orc.Simulacao(cena, iteracoes=op.get("iteracoes", 3), tensao=op.get("tensao", 0.0728), ar=op.get("ar", True),
                        mode=0)
    P0 = np.frombuffer(base64.b64decode(run["p0"]), np.float32)
    sim.adicionar(P0, False)
    assert sim.n == run["n0"]
    quadro = 1 / 60
    foam_diff = 0
    sub_index = 0
    eventos = [e for e in cfg.get("eventos", [])]
    n_ev = 0
    for fi, fr in enumerate(run["frames"]):
        mine = [e for e in eventos if e["quadro"] == fi]
        assert len(mine) == len(fr.get("eventos", [])), (nome, fi)
        n_ev += len(mine)
        for e, want in zip(mine, fr.get("eventos", [])):
            if e["tipo"] == "agua":
                r = e["retangulo"]
                lista = orc.semear_retangulo(r[0], r[1], r[2], r[3], h, e["absorcao"], orc.Semente(e["semente"]))
                assert sim.adicionar(lista, e["checar"]) == want["aceitas"], (nome, fi, "adicionar")
            else:
                t = e["ferramenta"]
                sim.ferramenta(t["tipo"], t.get("x", 0.0), t.get("y", 0.0), t.get("raio", 0.0), t.get("vx", 0.0),
                               t.get("vy", 0.0), t.get("absorcao", (0, 0, 0)))
            assert record(orc, sim)["s"] == want["s"], (nome, fi, e["tipo"])
        # the substep count the reference took from passoMaximo(vMax) must be ours too
        vmax = sim.vmax
        assert int(np.ceil(quadro / orc.passo_maximo(h, ny * h, vmax, quadro))) == fr["sub"], (nome, fi)
        if fr["ar"] is not None:
            sim.fase(orc.PH_AR)
            got = record(orc, sim)
            assert got["s"] == fr["ar"]["s"], (nome, "frame %d atualizarAr" % fi)
        d = quadro / fr["sub"]
        for si, ph in enumerate(fr["phases"]):
            for k, want in enumerate(ph):
                sim.fase(k, d)
                got = record(orc, sim)
                assert got["n"] == want["n"] and got["s"] == want["s"], (nome, "frame %d substep %d after %s" % (fi, si, PHASES[k]))
                foam_diff += got["f"] != want["f"]
            sub_index += 1
            cp = run["checkpoints"].get(str(sub_index))
            if cp is not None:
                want = np.frombuffer(base64.b64decode(cp), np.float32).reshape(-1, 12)
                P = sim.particulas()
                assert np.array_equal(np.delete(P, 8, axis=1), np.delete(want, 8, axis=1)), (nome, sub_index)
    assert n_ev == len(eventos)
    print("%s: %d substeps bit-identical; foam differed after %d phases" % (nome, sub_index, foam_diff))
    assert foam_diff == 0
