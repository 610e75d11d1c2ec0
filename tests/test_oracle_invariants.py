"""Pins of the oracle that do not come from the reference's own tests: closed forms, invariants
and independent solvers, each chosen so that a plausible restatement mistake (a dropped term, a
wrong sign or index, a transposed operand) fails it.

  operator A (aplicar, pressao-cpu.js:84-93)     symmetric on the water cells
  V-cycle (cicloV, pressao-cpu.js:241-255)       symmetric positive definite on the MG water cells
  PCG + bubble rows (pressao-cpu.js:271-402)     FP64 oracle solve = sparse direct solve of the
                                                 system written from the reference's equations
  tracker (RastreadorDeAr, ar.js:56-153)         labels = 4-connected components of a flood fill,
                                                 numbered by the reference's rules; split/merge
                                                 volumes follow the overlap shares
  APIC transfers (fisica-cpu.js:198-267, 556-661) reproduce an affine velocity field exactly
  surface (fisica-cpu.js:306-374)                disc of radius R: curvature 1/R on the interface
"""
import numpy as np
import pytest
from scipy import ndimage, sparse
from scipy.sparse.linalg import spsolve

from gpu_common import tanque_com_bolha

AGUA, AR, SOLIDO = 1, 2, 0
MOLDURA = 2


def _bubble_state(orc, frames=1):
    """FP64 oracle on the reference's bubble test scene, advanced to a state with water, a
    tracked bubble, atmosphere and walls; phases up to the bubble rhs run for the next solve."""
    osim = tanque_com_bolha(orc, mode=1)
    for _ in range(frames):
        osim.avancar(1 / 60, 4)
    dt = 1 / 240
    for ph in range(8):
        osim.fase(ph, dt)
    return osim, dt


def test_operator_is_symmetric(oracle):
    osim, _ = _bubble_state(oracle)
    tipo = osim.read(oracle.F_TIPO)
    water = tipo == AGUA
    rng = np.random.default_rng(1)
    x = np.where(water, rng.standard_normal(water.size), 0.0)
    y = np.where(water, rng.standard_normal(water.size), 0.0)
    ax, ay = osim.aplicar(x), osim.aplicar(y)
    lhs, rhs = y @ ax, x @ ay
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(y) * np.linalg.norm(ax)
    assert x @ ax > 0  # positive on a system with Dirichlet air


def test_vcycle_is_symmetric_positive_definite(oracle):
    osim, _ = _bubble_state(oracle)
    tipo = osim.read(oracle.F_TIPO)
    rot = osim.read(oracle.F_ROT)
    mgwater = (tipo == AGUA) | ((tipo == AR) & (rot >= 2))
    rng = np.random.default_rng(2)
    for _ in range(3):
        x = np.where(mgwater, rng.standard_normal(tipo.size), 0.0)
        y = np.where(mgwater, rng.standard_normal(tipo.size), 0.0)
        mx, my = osim.vcycle(x), osim.vcycle(y)
        assert abs(y @ mx - x @ my) <= 1e-11 * np.linalg.norm(y) * np.linalg.norm(mx)
        assert x @ mx > 0
        assert np.all(mx[~mgwater] == 0)
    # linear: M(2x + y) = 2 Mx + My
    assert np.allclose(osim.vcycle(2 * x + y), 2 * mx + my, rtol=1e-12, atol=1e-12 * np.abs(mx).max())


def test_pcg_with_bubble_matches_direct_solve(oracle):
    """Water rows: n q_c - sum(q of water neighbours) - sum(y of bubble-cell neighbours) = -div_c
    (walls excluded from n, other air is Dirichlet 0); bubble rows: sum over the bubble's interior
    cells' water faces of (y - q_w) = rhs_b."""
    osim, dt = _bubble_state(oracle)
    nx, ny = osim.nx, osim.ny
    tipo = osim.read(oracle.F_TIPO).astype(np.int64)
    rot = osim.read(oracle.F_ROT).astype(np.int64)
    div = osim.read(oracle.F_DIV, True)
    rhs_b = osim.read(oracle.F_RHSBOLHA, True)
    pairs = []
    for j in range(1, ny - 1):
        for i in range(1, nx - 1):
            c = i + j * nx
            if rot[c] < 2 or tipo[c] != AR:
                continue
            for d in (c - 1, c + 1, c - nx, c + nx):
                if tipo[d] == AGUA:
                    pairs.append((rot[c], d))
    active = sorted({b for b, _ in pairs})
    assert active, "the scene must hold a tracked bubble"
    widx = -np.ones(nx * ny, np.int64)
    wc = np.nonzero(tipo == AGUA)[0]
    widx[wc] = np.arange(len(wc))
    bidx = {b: len(wc) + k for k, b in enumerate(active)}
    n = len(wc) + len(active)
    rows, cols, vals = [], [], []
    rhs = np.zeros(n)
    for c in wc:
        i, j = c % nx, c // nx
        r = widx[c]
        cnt = 0
        for d, ok in ((c - 1, i > 0), (c + 1, i < nx - 1), (c - nx, j > 0), (c + nx, j < ny - 1)):
            if not ok or tipo[d] == SOLIDO:
                continue
            cnt += 1
            if tipo[d] == AGUA:
                rows.append(r); cols.append(widx[d]); vals.append(-1.0)
            elif rot[d] in bidx:
                rows.append(r); cols.append(bidx[rot[d]]); vals.append(-1.0)
        rows.append(r); cols.append(r); vals.append(float(cnt))
        rhs[r] = -div[c]
    for b, w in pairs:
        r = bidx[b]
        rows.append(r); cols.append(r); vals.append(1.0)
        rows.append(r); cols.append(widx[w]); vals.append(-1.0)
    for b in active:
        rhs[bidx[b]] = rhs_b[b]
    A = sparse.csr_matrix((vals, (rows, cols)), shape=(n, n))
    sol = spsolve(A, rhs)
    osim.set_tol(1e-13, 4000)
    osim.fase(oracle.PH_PRESSAO, dt)
    q = osim.read(oracle.F_Q, True)
    yb = osim.read(oracle.F_YBOLHA, True)
    st = osim.solver_stats()
    got = np.concatenate([q[wc], yb[active]])
    err = np.linalg.norm(got - sol) / np.linalg.norm(sol)
    print("\n  direct vs oracle PCG (%d water cells, bubbles %s, %d iterations): rel err %.2e" % (
        len(wc), active, st["iters"], err))
    assert err < 1e-9


def _tracker_sim(orc, nx, ny):
    cena = orc.Cena(nx, ny, 0.01)
    return orc.Simulacao(cena, mode=1, ar=True)


def _frame_types(nx, ny, air):
    t = np.where(air, AR, AGUA).astype(np.uint8)
    t[:MOLDURA, :] = SOLIDO
    t[-MOLDURA:, :] = SOLIDO
    t[:, :MOLDURA] = SOLIDO
    t[:, -MOLDURA:] = SOLIDO
    return t


def _air_update(orc, sim, tipo2d):
    sim.write(orc.F_TIPO, tipo2d.reshape(-1))
    sim.set_tipos_prontos(True)
    sim.fase(orc.PH_AR)
    return sim.read(orc.F_ROT).reshape(tipo2d.shape).astype(np.int64), sim.read(orc.F_VALVO, True)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_tracker_labels_are_flood_fill_components(oracle, seed):
    nx, ny = 48, 36
    rng = np.random.default_rng(seed)
    field = ndimage.gaussian_filter(rng.standard_normal((ny, nx)), 1.2)
    tipo = _frame_types(nx, ny, field > 0.05)
    sim = _tracker_sim(oracle, nx, ny)
    sim.reiniciar_ar()
    rot, vol = _air_update(oracle, sim, tipo)
    comp, ncomp = ndimage.label(tipo == AR, structure=[[0, 1, 0], [1, 1, 1], [0, 1, 0]])
    topo = ny - MOLDURA - 1
    # components in the order of their first cell in raster order (the union-find numbering)
    first = {}
    for c, k in enumerate(comp.reshape(-1)):
        if k and k not in first:
            first[k] = c
    order = sorted(first, key=lambda k: first[k])
    size = {k: int((comp == k).sum()) for k in order}
    top = {k: bool((comp[topo:, :] == k).any()) for k in order}
    cands = [k for k in order if not top[k] and size[k] >= 4]
    cands.sort(key=lambda k: -size[k])  # stable: equal sizes keep raster order
    expect = np.zeros((ny, nx), np.int64)
    for k in order:
        if top[k]:
            expect[comp == k] = 1
    for lab, k in enumerate(cands, start=2):
        expect[comp == k] = lab
    assert np.array_equal(rot, expect), "labels differ from the flood-fill numbering"
    for lab, k in enumerate(cands, start=2):
        assert vol[lab] == size[k]
    print("\n  seed %d: %d components, %d bubbles, %d atmosphere" % (seed, ncomp, len(cands), sum(top.values())))


def test_tracker_split_and_merge_share_volumes(oracle):
    nx, ny = 40, 30
    air = np.zeros((ny, nx), bool)
    air[8:14, 10:20] = True            # one 10x6 pocket (60 cells)
    sim = _tracker_sim(oracle, nx, ny)
    sim.reiniciar_ar()
    rot, vol = _air_update(oracle, sim, _frame_types(nx, ny, air))
    assert set(np.unique(rot)) == {0, 2} and vol[2] == 60
    split = air.copy()
    split[8:14, 14] = False            # a water column: left 4x6 = 24, right 5x6 = 30 cells
    rot, vol = _air_update(oracle, sim, _frame_types(nx, ny, split))
    left, right = rot[10, 11], rot[10, 16]
    # the larger share keeps the old label, the other takes the first free one
    assert right == 2 and left == 3
    assert vol[2] == pytest.approx(60 * 30 / 54, rel=1e-6)
    assert vol[3] == pytest.approx(60 * 24 / 54, rel=1e-6)
    rot, vol = _air_update(oracle, sim, _frame_types(nx, ny, air))
    assert set(np.unique(rot[air])) == {2}
    assert vol[2] == pytest.approx(60.0, rel=1e-6) and vol[3] == 0


def test_apic_transfers_reproduce_an_affine_field(oracle):
    nx, ny, h = 40, 30, 0.01
    cena = oracle.Cena(nx, ny, h)
    sim = oracle.Simulacao(cena, mode=1, tensao=0.0, ar=False)
    P = oracle.semear_retangulo(0.06, 0.06, 0.34, 0.24, h, (0, 0, 0), oracle.Semente(7)).reshape(-1, 12).copy()
    a, b, c, d, e, f = 0.3, 1.5, -0.7, -0.2, 0.4, 0.9
    x, y = P[:, 0].astype(np.float64), P[:, 1].astype(np.float64)
    P[:, 2] = a + b * x + c * y
    P[:, 3] = d + e * x + f * y
    P[:, 4], P[:, 5], P[:, 6], P[:, 7] = b, c, e, f
    sim.write(oracle.F_PART, P)
    sim.fase(oracle.PH_P2G)
    u = sim.read(oracle.F_U, True).reshape(ny, nx + 1)
    v = sim.read(oracle.F_V, True).reshape(ny + 1, nx)
    # faces well inside the seeded block (every weight of their stencil present)
    jj, ii = np.mgrid[9:22, 9:32]
    xu, yu = ii * h, (jj + 0.5) * h
    eu = np.abs(u[jj, ii] - (a + b * xu + c * yu)).max()
    xv, yv = (ii + 0.5) * h, jj * h
    ev = np.abs(v[jj, ii] - (d + e * xv + f * yv)).max()
    print("\n  P2G affine: max |u - u_exact| %.2e, |v - v_exact| %.2e" % (eu, ev))
    # the particle record is binary32 (velocities round at ~6e-8 relative); a missing or
    # mis-signed APIC term would err by ~|grad| * h/2 ~ 1e-2
    assert eu < 2e-7 and ev < 2e-7
    # G2P gathers the same field back (velocity and its gradient) before moving the particles
    sim.fase(oracle.PH_G2P, 1e-6)
    Q = sim.read(oracle.F_PART, True).reshape(-1, 12)
    inner = (x > 0.1) & (x < 0.3) & (y > 0.1) & (y < 0.2)
    evx = np.abs(Q[inner, 2] - (a + b * x[inner] + c * y[inner])).max()
    evy = np.abs(Q[inner, 3] - (d + e * x[inner] + f * y[inner])).max()
    ec = np.abs(Q[inner, 4:8] - np.array([b, c, e, f])).max()
    print("  G2P affine: max |v - v_exact| %.2e %.2e, |C - grad| %.2e (%d particles)" % (evx, evy, ec, inner.sum()))
    assert evx < 2e-7 and evy < 2e-7 and ec < 1e-4


@pytest.mark.parametrize("R", [0.06, 0.10])
def test_disc_curvature_is_one_over_radius(oracle, R):
    nx, ny, h = 60, 40, 0.01
    cena = oracle.Cena(nx, ny, h)
    sim = oracle.Simulacao(cena, mode=1)
    m = MOLDURA * h
    P = oracle.semear_retangulo(m, m, nx * h - m, ny * h - m, h, (0, 0, 0), oracle.Semente(4)).reshape(-1, 12)
    x, y = P[:, 0].astype(np.float64), P[:, 1].astype(np.float64)
    sim.adicionar(P[(x - 0.3) ** 2 + (y - 0.2) ** 2 < R * R].reshape(-1), False)
    for ph in (oracle.PH_P2G, oracle.PH_TIPOS, oracle.PH_SUPERFICIE):
        sim.fase(ph, 1 / 240)
    k = sim.read(oracle.F_KAPPA, True)
    frac = sim.read(oracle.F_FASE, True)
    iface = (frac > 0.2) & (frac < 0.8)
    mean = k[iface].mean()
    print("\n  disc R=%.2f m: mean curvature on %d interface cells %.3f 1/m (1/R = %.3f)" % (R, iface.sum(), mean, 1 / R))
    assert abs(mean - 1 / R) < 0.05 / R
