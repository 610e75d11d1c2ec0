"""Shared helpers of the GPU parity tests: scenario builders on the oracle, cloning the oracle's
state into the HIP simulation, and comparison helpers that print what they compare."""
import math
import sys

import numpy as np

from modelos import MODELOS, montar_modelo, semear_agua

STATE_FIELDS = list(range(0, 25))  # every field id except particles
MOLDURA = 2


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def clone_state(osim, gsim):
    """Copies the oracle's complete state (particles, grid, labels, bubbles) into the HIP sim."""
    gsim.write_particles(osim.particulas())
    for f in STATE_FIELDS:
        if f == 10:  # solid cells: derived from the same SDF on both sides
            continue
        gsim.write(f, osim.read(f))
    gsim.set_primeira(bool(osim_primeira(osim)))


def osim_snapshot(osim):
    """The oracle's complete state as arrays (particles + every field + tracker flag)."""
    return {"P": osim.particulas().copy(), "fields": {f: osim.read(f) for f in STATE_FIELDS},
            "primeira": osim_primeira(osim), "cena": osim.cena,
            "opts": getattr(osim, "_opts", None)}


def oracle_fp64_clone(orc, snap, **opts):
    """An FP64-storage oracle holding the snapshot's state (binary32 values are exact in FP64)."""
    o = orc.Simulacao(snap["cena"], mode=1, **opts)
    o.write(orc.F_PART, snap["P"])
    for f, arr in snap["fields"].items():
        o.write(f, arr)
    o.set_primeira(bool(snap["primeira"]))
    return o


def osim_primeira(osim):
    from oracle import oracle as orc
    return orc.lib().or_sim_tracker_first(osim._p)


def hip_sim_like(tq, osim, **kw):
    opts = dict(gravidade=kw.pop("gravidade", (0.0, -9.81)), iteracoes=kw.pop("iteracoes", 3),
                tensao=kw.pop("tensao", 0.0728), ar=kw.pop("ar", True))
    opts.update(kw)
    return tq.SimulacaoHIP(osim.cena, **opts)


def tanque_com_bolha(orc, ar=True, mode=0):
    """testes/fisica.test.js:163-177: 60x60 tank, water up to 0.45 m with a 5 cm air hole."""
    h = 0.01
    cena = orc.Cena(60, 60, h)
    sim = orc.Simulacao(cena, ar=ar, mode=mode)
    m = MOLDURA * h
    agua = orc.semear_retangulo(m, m, 0.6 - m, 0.45, h, (0, 0, 0), orc.Semente(5)).reshape(-1, 12)
    x, y = agua[:, 0].astype(np.float64), agua[:, 1].astype(np.float64)
    fora = agua[~((x - 0.3) ** 2 + (y - 0.12) ** 2 < 0.05 ** 2)]
    sim.adicionar(fora.reshape(-1), False)
    return sim


def copo(orc, mode=0):
    _, sim, _ = montar_modelo(orc, "copo", 90, 60, 0.01, orc.Semente(11), mode=mode)
    return sim


def ampulheta(orc, ar=True, mode=0):
    h = 0.01
    cena = orc.Cena(60, 90, h)
    m = MODELOS["ampulheta"](cena.largura, cena.altura, h)
    for t in m["tracos"]:
        cena.tracar(*t)
    sim = orc.Simulacao(cena, ar=ar, mode=mode)
    for a in m["agua"]:
        sim.adicionar(semear_agua(orc, a, h, (0, 0, 0), orc.Semente(2)), False)
    return sim


def dam(orc, nx=80, ny=60, h=0.01, mode=0, seed=9):
    cena = orc.Cena(nx, ny, h)
    sim = orc.Simulacao(cena, mode=mode)
    m = MOLDURA * h
    sim.adicionar(orc.semear_retangulo(m, m, 0.3, 0.42, h, (0, 0, 0), orc.Semente(seed)), False)
    return sim


def substep_dt(orc, osim, quadro=1 / 60):
    sub = math.ceil(quadro / orc.passo_maximo(osim.h, osim.cena.altura, osim.vmax, quadro))
    return quadro / sub


def cmp_exact(name, a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (name, a.shape, b.shape)
    if a.dtype.kind == "f":
        same = (a.view(np.uint32 if a.dtype == np.float32 else np.uint64) ==
                b.view(np.uint32 if b.dtype == np.float32 else np.uint64)) | (np.isnan(a) & np.isnan(b))
    else:
        same = a == b
    bad = int((~same).sum())
    extra = ""
    if bad and a.dtype.kind == "f":
        d = np.abs(a.astype(np.float64) - b.astype(np.float64))
        extra = " max|d|=%.3g at %d" % (np.nanmax(d), int(np.nanargmax(d)))
    elif bad:
        k = int(np.nonzero(~same)[0][0])
        extra = " first at %d: %s vs %s" % (k, a[k], b[k])
    log("  exact %-14s n=%-7d mismatches=%d%s" % (name, a.size, bad, extra))
    return bad


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.max(np.abs(b)) if b.size else 0.0, 1e-30)
    return float(np.max(np.abs(a - b)) / scale) if a.size else 0.0


def cmp_close(name, a, b, rtol):
    e = rel_err(a, b)
    log("  close %-14s n=%-7d max|d|/max|ref|=%.3e (tol %.1e)" % (name, np.asarray(a).size, e, rtol))
    return e <= rtol


def tanque_com_bolhas(orc, n=12, mode=0):
    """A 120x80 tank of water (up to 0.62 m) with n enclosed round air holes of radius 2.5 cm
    on a grid, more bubbles than the solver's per-block slots."""
    h = 0.01
    cena = orc.Cena(120, 80, h)
    sim = orc.Simulacao(cena, mode=mode)
    m = MOLDURA * h
    agua = orc.semear_retangulo(m, m, 1.2 - m, 0.62, h, (0, 0, 0), orc.Semente(13)).reshape(-1, 12)
    x, y = agua[:, 0].astype(np.float64), agua[:, 1].astype(np.float64)
    keep = np.ones(len(agua), bool)
    cols = 6
    for k in range(n):
        cx = 0.12 + 0.19 * (k % cols)
        cy = 0.14 + 0.2 * (k // cols)
        keep &= ~((x - cx) ** 2 + (y - cy) ** 2 < 0.025 ** 2)
    sim.adicionar(np.ascontiguousarray(agua[keep]).reshape(-1), False)
    return sim
