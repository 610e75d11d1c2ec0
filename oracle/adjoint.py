"""ctypes binding of the FP64 replay and its adjoint (oracle/adjoint_oracle.cpp), and the tape
recorder that drives the JS-exact oracle (oracle/tanque_oracle.cpp) phase by phase.

TEST INFRASTRUCTURE ONLY: the checker the HIP adjoint is compared against.  Only tests/ and
bench.py may import this module.

A tape holds every discrete decision of a run (see adjoint_oracle.cpp): per frame the emitted
particles and the substep count; per substep the cell types, the active bubble of each air
cell, the air pressures of surface tension, the frozen pressure of bubbles without water
faces, the frozen part of each bubble's right-hand side; per particle and substep the flag word
(micro-step count, collision steps, velocity projections, removal).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_adj.so")
SRC = os.path.join(HERE, "adjoint_oracle.cpp")

AR, AGUA, SOLIDO = 2, 1, 0
VOLTA_VOLUME_BOLHA = 0.05   # constantes.js:29
DENSIDADE_AGUA = 1000.0     # constantes.js:11

_lib = None


def build(force=False):
    import subprocess
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(SRC):
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                               "-fPIC", "-shared", SRC, "-o", LIB_PATH + ".tmp"])
        os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, d, i = C.c_void_p, C.c_double, C.c_int
        L.adj_run.restype = i
        L.adj_run.argtypes = [i, i, d, d, d, d, vp, i, vp, i, vp, vp, vp, vp, vp, vp, vp, d, vp, vp, vp,
                              i, C.c_long, vp, vp, vp, d, d, vp, i, vp, vp, vp]
        L.adj_apply_operator.restype = None
        L.adj_apply_operator.argtypes = [i, i, vp, vp, i, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data if a is not None else None


def bubble_fields(nx, ny, h, dt, tipo, rot, qar, vol_alvo, y_bolha):
    """The bubble part of a substep's tape from the fields the reference holds after
    #ladoDireitoDasBolhas (fisica-cpu.js:430-456) and before #pressao (:487-504): the active
    bubbles (labels with a water face, in first-seen raster order as montarBolhas,
    pressao-cpu.js:271-288), bub[c] for their air cells, qextra (the warm-start y of bubbles
    with cells but no water face: #pressao adds it to their cells' q) and rhsfix, the volume
    return of each active bubble's rhs (the face velocities and air pressures are the replay's
    own).  qar is unused (the replay computes the air pressures)."""
    C_ = nx * ny
    tipo = np.asarray(tipo).reshape(-1)
    rot = np.asarray(rot).reshape(-1).astype(np.int64)
    qar = np.asarray(qar, np.float64).reshape(-1)
    bub = np.full(C_, -1, np.int32)
    qextra = np.zeros(C_)
    order, idx = [], {}
    vol = {}
    for j in range(1, ny - 1):
        for i in range(1, nx - 1):
            c = i + j * nx
            r = int(rot[c])
            if tipo[c] != AR or r < 2:
                continue
            vol[r] = vol.get(r, 0) + 1
            for nbh in (c - 1, c + 1, c - nx, c + nx):
                if tipo[nbh] == AGUA and r not in idx:
                    idx[r] = len(order)
                    order.append(r)
    rhsfix = np.zeros(len(order))
    for r, b in idx.items():
        rhsfix[b] = VOLTA_VOLUME_BOLHA * (float(vol_alvo[r]) - vol[r]) * h / dt
    air = (tipo == AR) & (rot >= 2)
    for c in np.nonzero(air)[0]:
        r = int(rot[c])
        if r in idx:
            bub[c] = idx[r]
        elif vol.get(r, 0) > 0:
            qextra[c] = float(y_bolha[r])
    return bub, qextra, rhsfix


class Tape:
    """Tape of one run (arrays as adj_run takes them)."""

    def __init__(self, nx, ny, h, gx=0.0, gy=-9.81, tau=0.1, sigma=0.0728):
        self.nx, self.ny, self.h, self.gx, self.gy, self.tau, self.sigma = nx, ny, h, gx, gy, tau, sigma
        self.f_nemit, self.f_nsub, self.f_dt, self.emit_xy = [], [], [], []
        self.tipo, self.bub, self.qar, self.qextra, self.nb, self.rhsfix = [], [], [], [], [], []
        self.np = None
        self.flags = None

    def add_frame(self, emit_xy, nsub, dt):
        self.f_nemit.append(len(emit_xy))
        self.f_nsub.append(nsub)
        self.f_dt.append(dt)
        self.emit_xy.extend([float(v) for xy in emit_xy for v in xy])

    def add_substep(self, tipo, bub, qar, qextra, rhsfix):
        self.tipo.append(np.asarray(tipo, np.uint8).reshape(-1))
        self.bub.append(np.asarray(bub, np.int32).reshape(-1))
        self.qar.append(np.asarray(qar, np.float64).reshape(-1))
        self.qextra.append(np.asarray(qextra, np.float64).reshape(-1))
        self.nb.append(len(rhsfix))
        self.rhsfix.extend([float(v) for v in rhsfix])

    def arrays(self):
        a = {
            "f_nemit": np.asarray(self.f_nemit, np.int32), "f_nsub": np.asarray(self.f_nsub, np.int32),
            "f_dt": np.asarray(self.f_dt, np.float64), "emit_xy": np.asarray(self.emit_xy + [0.0], np.float64),
            "tipo": np.concatenate(self.tipo), "bub": np.concatenate(self.bub),
            "qar": np.concatenate(self.qar), "qextra": np.concatenate(self.qextra),
            "nb": np.asarray(self.nb, np.int32), "rhsfix": np.asarray(self.rhsfix + [0.0], np.float64),
        }
        return a


def run(tape, sdf, p0, vjet, box, target=0.5, record=False, grad=True, cg_tol=1e-13, flags_cap=None):
    """The replay over the tape, and its adjoint: returns dict(loss, frac, g_vjet, g_sdf, g_p0).
    record=True decides the per-particle flags (stored in the tape)."""
    a = tape.arrays()
    S = int(a["f_nsub"].sum())
    p0 = np.ascontiguousarray(np.asarray(p0, np.float64).reshape(-1, 8))
    sdf = np.ascontiguousarray(np.asarray(sdf, np.float64).reshape(-1))
    vjet = np.ascontiguousarray(np.asarray(vjet, np.float64))
    box = np.ascontiguousarray(np.asarray(box, np.int32))
    if record:
        cap = flags_cap or int((len(p0) + len(a["emit_xy"])) * S + 16)
        np_ = np.zeros(S, np.int32)
        flags = np.zeros(cap, np.uint16)
    else:
        np_, flags = tape.np, tape.flags
        cap = len(flags)
    out = np.zeros(2)
    g_vf = np.zeros(2)
    g_sdf = np.zeros(len(sdf)) if grad else None
    g_p0 = np.zeros(max(1, p0.size)) if grad else None
    rc = lib().adj_run(tape.nx, tape.ny, tape.h, tape.gx, tape.gy, tape.tau, _p(sdf), len(p0), _p(p0),
                       len(a["f_nsub"]), _p(a["f_nemit"]), _p(a["f_nsub"]), _p(a["f_dt"]), _p(a["emit_xy"]),
                       _p(vjet), _p(a["tipo"]), _p(a["bub"]), tape.sigma, _p(a["qextra"]), _p(a["nb"]),
                       _p(a["rhsfix"]), 1 if record else 0, cap, _p(np_), _p(flags), _p(box), target, cg_tol,
                       _p(out), 1 if grad else 0, _p(g_vf), _p(g_sdf), _p(g_p0))
    if rc != 0:
        raise RuntimeError("adj_run failed: %d" % rc)
    if record:
        tape.np = np_
        tape.flags = flags[: int(np_.astype(np.int64).sum())].copy()
    return {"loss": out[0], "frac": out[1], "g_vjet": g_vf, "g_sdf": g_sdf,
            "g_p0": g_p0[: p0.size].reshape(-1, 8) if grad else None}


def apply_operator(nx, ny, tipo, bub, nb, z):
    z = np.ascontiguousarray(np.asarray(z, np.float64))
    out = np.zeros_like(z)
    lib().adj_apply_operator(nx, ny, _p(np.ascontiguousarray(tipo, dtype=np.uint8)),
                             _p(np.ascontiguousarray(bub, dtype=np.int32)), nb, _p(z), _p(out))
    return out


def record_with_oracle(work, frames=None, mode=0, sub_cap=None):
    """Runs the JS-exact oracle (mode 0) over `work` (a scenes.Workload with a faucet) frame by
    frame as the app does (faucet, then avancar with the substeps of passoMaximo), phase by phase,
    and records the tape's cell fields; returns (tape, sdf, p0, vjet).  The per-particle flags
    are recorded afterwards by run(record=True) on the replay's own trajectory."""
    from oracle import oracle as orc
    nx, ny, h = work.nx, work.ny, work.h
    c = orc.Cena(nx, ny, h)
    for s in work.strokes:
        c.tracar(*s)
    sim = orc.Simulacao(c, mode=mode)
    P = np.asarray(work.particles, np.float32).reshape(-1, 12)
    if len(P):
        sim.adicionar(P.reshape(-1), False)
    p0 = np.zeros((len(P), 8))
    p0[:, :] = P[:, :8]
    f = work.extra["faucet"]
    sem = orc.Semente(f["seed"])
    pvx, pvy = f["pvx"], f["pvy"]
    vjet = np.array([0.5 * pvx, 0.5 * pvy - 0.7])
    tape = Tape(nx, ny, h)
    frames = work.frames if frames is None else frames
    dt = work.dt_frame
    H = ny * h
    for k in range(frames):
        n0 = sim.n
        emitted = []
        if k < f["frames"]:
            sim.emitir_torneira(f["x"], f["y"], pvx, pvy, f["raio"], f["absorcao"], dt, sem)
            Pn = sim.particulas()
            emitted = [(float(Pn[i, 0]), float(Pn[i, 1])) for i in range(n0, sim.n)]
        vmax = sim.vmax
        nsub = int(np.ceil(dt / orc.passo_maximo(h, H, vmax, dt)))
        if sub_cap:
            nsub = min(nsub, sub_cap)
        tape.add_frame(emitted, nsub, dt)
        sim.fase(orc.PH_AR)
        d = dt / nsub
        for _ in range(nsub):
            for ph in (orc.PH_P2G, orc.PH_GRAV, orc.PH_TIPOS, orc.PH_ROTULOS, orc.PH_SUPERFICIE, orc.PH_PAR,
                       orc.PH_CLASSIFICAR, orc.PH_BOLHAS):
                sim.fase(ph, d)
            tipo = sim.read(orc.F_TIPO)
            rot = sim.read(orc.F_ROT)
            qar = sim.read(orc.F_QAR, as_double=True)
            valvo = sim.read(orc.F_VALVO, as_double=True)
            yb = sim.read(orc.F_YBOLHA, as_double=True)
            bub, qextra, rhsfix = bubble_fields(nx, ny, h, d, tipo, rot, qar, valvo, yb)
            tape.add_substep(tipo, bub, qar, qextra, rhsfix)
            for ph in (orc.PH_PRESSAO, orc.PH_GRADIENTE, orc.PH_EXTRAPOLAR, orc.PH_G2P):
                sim.fase(ph, d)
    return tape, np.asarray(c.sdf, np.float64), p0, vjet, sim
