"""Synthetic workloads of the benchmark configurations, built as capsule strokes (the
reference's Cena semantics) and seeded particle rectangles.

Each builder returns a `Workload` with the grid, the strokes (metres), the particle list (the
reference's 12-float records, in insertion order), the frame dt and a description; the same
strokes and particles drive the HIP path and the CPU baseline, so both see identical inputs.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import Cena, Semente, semear_retangulo

TANK_W, TANK_H = 1.42, 0.80  # metres
FRAME = 2                    # frame thickness in cells (MOLDURA)


@dataclass
class Workload:
    name: str
    nx: int
    ny: int
    h: float
    strokes: list
    particles: np.ndarray
    dt_frame: float
    frames: int
    description: str
    extra: dict = field(default_factory=dict)

    def cena(self):
        c = Cena(self.nx, self.ny, self.h)
        for s in self.strokes:
            c.tracar(*s)
        return c


def dam(seed=0):
    """Plain dam break: 160x90 grid, 8.9 mm cells, empty tank (2-cell frame, no lid); a water
    block x < 40 %, y < 50 % of the tank at 4 particles/cell (seeded jitter); 240 frames at
    1/240 s."""
    nx, ny = 160, 90
    h = TANK_H / ny
    W, H = nx * h, ny * h
    m = FRAME * h
    P = semear_retangulo(m, m, 0.4 * W, 0.5 * H, h, (0, 0, 0), Semente(seed))
    return Workload("dam", nx, ny, h, [], P, 1 / 240, 240,
                    "dam break 160x90 h=%.4g m, block 40%%x50%% at 4 ppc" % h)


def hourglass_strokes(nx, ny, h, neck_cells=10, half_height=0.12, radius_cells=3):
    """Two solid wedges from the side walls to the middle, leaving `neck_cells` between their
    tips at mid-height: each wedge is a fan of capsules from the wall to its tip."""
    W, H = nx * h, ny * h
    r = radius_cells * h
    tip_l = W / 2 - (neck_cells / 2) * h - r
    tip_r = W / 2 + (neck_cells / 2) * h + r
    n = int(math.ceil(2 * half_height / r))
    strokes = []
    for k in range(n + 1):
        y = H / 2 - half_height + k * (2 * half_height / n)
        strokes.append((0.0, y, tip_l, H / 2, r, False))
        strokes.append((W, y, tip_r, H / 2, r, False))
    return strokes


def ultra_hourglass(seed=0, nx=896, ny=504):
    """Ultra ~450k cells: 896x504 (1.59 mm cells), sealed hourglass -- two solid wedges leaving a
    10-cell neck, closed by the tank's top frame; the upper chamber full of water at 4 ppc, the
    lower chamber one enclosed air pocket (a bubble unknown after the first air update);
    frames of 1/480 s."""
    h = TANK_H / ny
    W, H = nx * h, ny * h
    strokes = hourglass_strokes(nx, ny, h)
    m = FRAME * h
    P = semear_retangulo(m, H / 2, W - m, H - m, h, (0, 0, 0), Semente(seed))
    return Workload("ultra", nx, ny, h, strokes, P, 1 / 480, 480,
                    "sealed hourglass %dx%d h=%.4g m, 10-cell neck, upper chamber full at 4 ppc" % (nx, ny, h))


def alto_cup(seed=0, nx=480, ny=270):
    """Alto ~130k cells: 480x270 (2.96 mm cells), a U-shaped cup (three strokes of radius 3h: 6
    cells thick, 120 cells between the inner faces, centred, from the floor to 46 % of the
    stage height) and a faucet above it: the faucet tool (emitirTorneira) of radius 6h (12 cells)
    at the top, the pointer moving down at 1.6 m/s (water at 1.5 m/s), seeded with `seed`, on
    for the first second; 480 frames of 1/240 s.  No particles at the start."""
    h = TANK_H / ny
    W, H = nx * h, ny * h
    r = 3 * h
    m = FRAME * h
    cx = W / 2
    half = 60 * h + r                      # stroke centre line: inner face 60 cells from the centre
    base = m + r
    top = m + 0.46 * (H - 2 * m)
    strokes = [(cx - half, top, cx - half, base, r, False), (cx - half, base, cx + half, base, r, False),
               (cx + half, base, cx + half, top, r, False)]
    raio = 6 * h
    faucet = {"x": cx, "y": H - m - raio - 2 * h, "pvx": 0.0, "pvy": -1.6, "raio": raio,
              "absorcao": (0.0, 0.0, 0.0), "seed": seed, "frames": 240}
    w = Workload("alto", nx, ny, h, strokes, np.zeros(0, np.float32), 1 / 240, 480,
                 "cup 480x270 h=%.4g m (U, 6 cells thick, 120 inner), faucet 12 cells at 1.5 m/s for 1 s" % h,
                 extra={"faucet": faucet})
    return w


def maze_strokes(nx, ny, h, aleatorio, lattice=8, wall_cells=3):
    """The maze of modelos.js:133-185 (recursive backtracker from the top-left cell, entrance at
    the top of the first column and exit at the bottom of the last) on a lattice of `lattice`
    cells filling the tank inside its frame, walls `wall_cells` thick."""
    W, H = nx * h, ny * h
    tam = lattice * h
    cols = max(2, (nx - 2 * FRAME) // lattice)
    rows = max(2, (ny - 2 * FRAME) // lattice)
    x0, y0 = (W - cols * tam) / 2, (H - rows * tam) / 2
    r = wall_cells * h / 2
    dir_ = [[True] * rows for _ in range(cols)]
    baixo = [[True] * rows for _ in range(cols)]
    visto = [[False] * rows for _ in range(cols)]
    pilha = [(0, rows - 1)]
    visto[0][rows - 1] = True
    while pilha:
        c, l = pilha[-1]
        viz = [(a, b) for a, b in ((c + 1, l), (c - 1, l), (c, l + 1), (c, l - 1))
               if 0 <= a < cols and 0 <= b < rows and not visto[a][b]]
        if not viz:
            pilha.pop()
            continue
        a, b = viz[int(math.floor(aleatorio() * len(viz)))]
        if a == c + 1:
            dir_[c][l] = False
        elif a == c - 1:
            dir_[a][l] = False
        elif b == l + 1:
            baixo[c][b] = False
        else:
            baixo[c][l] = False
        visto[a][b] = True
        pilha.append((a, b))
    X = lambda c: x0 + c * tam
    Y = lambda l: y0 + l * tam
    t = []
    for c in range(cols):
        for l in range(rows):
            if dir_[c][l] and c < cols - 1:
                t.append((X(c + 1), Y(l), X(c + 1), Y(l + 1), r, False))
            if baixo[c][l] and l > 0:
                t.append((X(c), Y(l), X(c + 1), Y(l), r, False))
    t += [(X(0), Y(0), X(0), Y(rows), r, False), (X(cols), Y(0), X(cols), Y(rows), r, False),
          (X(1), Y(rows), X(cols), Y(rows), r, False), (X(0), Y(0), X(cols - 1), Y(0), r, False)]
    return t


def batched_mazes(n=64, seed=0, h=TANK_H / 270):
    """The batched configuration: n tanks with widths uniform in [64, 384] and heights in [48, 256]
    cells (h = 2.963 mm), each a maze (maze_strokes) and a water block over the full width and a
    uniform 30-60 % of the height at 4 ppc; all draws from the seeded LCG; frames of 1/240 s."""
    from . import Semente as _Semente
    rng = _Semente(seed)

    def u():
        rng.state.value = (rng.state.value * 1664525 + 1013904223) & 0xFFFFFFFF
        return rng.state.value / 4294967296.0

    tanks = []
    for k in range(n):
        nx = 64 + int(math.floor(u() * 321))
        ny = 48 + int(math.floor(u() * 209))
        strokes = maze_strokes(nx, ny, h, u)
        frac = 0.3 + 0.3 * u()
        W, H = nx * h, ny * h
        m = FRAME * h
        P = semear_retangulo(m, m, W - m, m + frac * (H - 2 * m), h, (0, 0, 0), _Semente(seed * 1000 + k + 1))
        tanks.append(Workload("maze%d" % k, nx, ny, h, strokes, P, 1 / 240, 500,
                              "maze %dx%d, water %.0f %% of the height" % (nx, ny, 100 * frac)))
    return tanks


def scaled_hourglass(nx, ny, seed=0):
    """The Ultra scene at another resolution (size sweeps past the 256 MB last-level cache)."""
    w = ultra_hourglass(seed, nx, ny)
    w.name = "hourglass%dx%d" % (nx, ny)
    return w


def passo_frame(h, altura, vmax, dt_frame):
    from . import passo_maximo
    return int(math.ceil(dt_frame / passo_maximo(h, altura, vmax, dt_frame)))


# ------------------------------------------------------------------ the funnel of the adjoint workload
FUNNEL_SEGMENTS = 16  # capsule segments per wall; one height parameter each


def funnel_geometry(nx=256, ny=144, segments=FUNNEL_SEGMENTS):
    """Centre lines of the two funnel walls (metres): the left wall from (0.38 W, 0.80 H) and the
    right from (0.93 W, 0.80 H) down to a spout at x = 0.68 W, y = 0.40 H whose inner faces are
    12 cells apart; walls 4 cells thick (capsules of radius 2h)."""
    h = TANK_H / ny
    W, H = nx * h, ny * h
    r = 2 * h
    xs, ys = 0.68 * W, 0.40 * H
    left = ((0.38 * W, 0.80 * H), (xs - 6 * h - r, ys))
    right = ((0.93 * W, 0.80 * H), (xs + 6 * h + r, ys))
    return h, r, [left, right], segments


def funnel_strokes(heights, nx=256, ny=144, segments=FUNNEL_SEGMENTS):
    """The funnel's capsule strokes: each wall split into `segments` equal pieces, piece k of
    wall w moved up by heights[w * segments + k] metres (both of its ends)."""
    h, r, walls, S = funnel_geometry(nx, ny, segments)
    heights = np.asarray(heights, np.float64)
    assert heights.shape == (2 * S,)
    strokes = []
    for w, (a, b) in enumerate(walls):
        for k in range(S):
            t0, t1 = k / S, (k + 1) / S
            dy = float(heights[w * S + k])
            strokes.append((a[0] + (b[0] - a[0]) * t0, a[1] + (b[1] - a[1]) * t0 + dy,
                            a[0] + (b[0] - a[0]) * t1, a[1] + (b[1] - a[1]) * t1 + dy, r, False))
    return strokes


def funnel(heights=None, vjet=(0.6, -1.5), seed=0, nx=256, ny=144, frames=120):
    """Adjoint 256x144 (5.56 mm cells): the funnel of `funnel_strokes`, and a faucet of radius 6h
    (a 12-cell jet) above the left half of the funnel emitting water at the jet velocity `vjet`
    (m/s) every frame (the faucet tool with the pointer velocity that gives that jet:
    principal.js:302 maps pointer velocity p to 0.5 p + (0, -0.7)), seeded with `seed`;
    `frames` frames of 1/240 s.  The loss box is the 40x30 cells at the bottom right inside
    the frame; the loss is (water fraction of the box - 0.5)^2."""
    h = TANK_H / ny
    W, H = nx * h, ny * h
    if heights is None:
        heights = np.zeros(2 * FUNNEL_SEGMENTS)
    strokes = funnel_strokes(heights, nx, ny)
    raio = 6 * h
    vx, vy = float(vjet[0]), float(vjet[1])
    faucet = {"x": 0.60 * W, "y": 0.90 * H, "pvx": 2 * vx, "pvy": 2 * (vy + 0.7), "raio": raio,
              "absorcao": (0.0, 0.0, 0.0), "seed": seed, "frames": frames}
    box = (nx - FRAME - 40, FRAME, nx - FRAME, FRAME + 30)
    return Workload("funnel", nx, ny, h, strokes, np.zeros(0, np.float32), 1 / 240, frames,
                    "funnel 256x144 h=%.4g m, 2x16 wall segments, faucet jet (%.2f, %.2f) m/s" % (h, vx, vy),
                    extra={"faucet": faucet, "box": box, "target": 0.5, "heights": np.asarray(heights, np.float64)})


def stroke_shift_vjp(nx, ny, h, strokes, gsdf, movable):
    """Chain rule from the SDF lattice to vertical shifts of strokes: returns, for each stroke
    index in `movable`, sum over lattice points of gsdf * d(sdf)/d(shift), for an SDF built as
    Cena.tracar builds it (js/cena.js:43-72: the frame, then the running minimum of clamped
    capsule distances, stored in binary32 after each stroke).  The point's value comes from the
    last stroke that lowered it (none: the frame, no derivative); its distance d = (|p| - rp) / r
    (r = 2 points per cell) moves with a shift s of both ends of that stroke by
    (p_y / |p|) / h per metre, p the vector from the point to its nearest point on the segment,
    and not at all where d is clamped to +-SDF_MAX."""
    R = 2
    mw, mh = nx * R, ny * R
    gsdf = np.asarray(gsdf, np.float64).reshape(mh, mw)
    X, Y = np.meshgrid(np.arange(mw, dtype=np.float64), np.arange(mh, dtype=np.float64))
    m = FRAME * R - 0.5
    cur = np.minimum(np.minimum(X - m, mw - 1 - m - X), np.minimum(Y - m, mh - 1 - m - Y)) / R
    cur = np.clip(cur, -10, 10).astype(np.float32)
    owner = np.full((mh, mw), -1, np.int64)
    deriv = np.zeros((mh, mw))
    s = R / h
    for k, (x0, y0, x1, y1, raio, apagar) in enumerate(strokes):
        assert not apagar
        ax, ay, bx, by = x0 * s - 0.5, y0 * s - 0.5, x1 * s - 0.5, y1 * s - 0.5
        rp = raio * s
        alc = rp + 10 * R + 1
        xa = max(0, int(math.floor(min(ax, bx) - alc))); xb = min(mw - 1, int(math.ceil(max(ax, bx) + alc)))
        ya = max(0, int(math.floor(min(ay, by) - alc))); yb = min(mh - 1, int(math.ceil(max(ay, by) + alc)))
        if xa > xb or ya > yb:
            continue
        px_, py_ = X[ya:yb + 1, xa:xb + 1], Y[ya:yb + 1, xa:xb + 1]
        dx, dy = bx - ax, by - ay
        len2 = dx * dx + dy * dy
        t = ((px_ - ax) * dx + (py_ - ay) * dy) / len2 if len2 > 0 else np.zeros_like(px_)
        t = np.clip(t, 0, 1)
        qx, qy = ax + t * dx - px_, ay + t * dy - py_
        nrm = np.sqrt(qx * qx + qy * qy)
        d = (nrm - rp) / R
        dc = np.clip(d, -10, 10)
        c = cur[ya:yb + 1, xa:xb + 1]
        new = np.minimum(c.astype(np.float64), dc).astype(np.float32)
        low = new < c
        with np.errstate(invalid="ignore", divide="ignore"):
            dd = np.where((np.abs(d) < 10) & (nrm > 0), qy / np.where(nrm > 0, nrm, 1) / h, 0.0)
        o = owner[ya:yb + 1, xa:xb + 1]
        o[low] = k
        dv = deriv[ya:yb + 1, xa:xb + 1]
        dv[low] = dd[low]
        cur[ya:yb + 1, xa:xb + 1] = new
    out = np.zeros(len(movable))
    for j, k in enumerate(movable):
        sel = owner == k
        out[j] = float(np.sum(gsdf[sel] * deriv[sel]))
    return out
