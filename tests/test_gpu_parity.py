"""Phase-by-phase parity of the HIP substep against the oracle on identical inputs.

For each scenario the oracle (JS-exact mode) is advanced to a mid-run state; then, phase by
phase through one substep, the oracle's full state is cloned into the HIP simulation, the same
phase runs on both, and the outputs are compared:
  * bit-exact: cell counts and types (P2G binning), labels (propagation, tracker), the surface
    chain, air pressure, divergence, gradient, extrapolation and G2P (given the same grid);
  * tolerance: the P2G sums (fixed-point exact on the GPU vs sequential binary32 in the
    reference), the per-bubble rhs (FP64 vs binary32 accumulation) and the 3-iteration
    warm-started PCG (different rounding inside the V-cycle).
"""
import numpy as np
import pytest

from gpu_common import (ampulheta, clone_state, cmp_close, cmp_exact, copo, dam, hip_sim_like, log,
                        oracle_fp64_clone, osim_snapshot, rel_err, substep_dt, tanque_com_bolha)

pytestmark = pytest.mark.gpu

# tolerances (relative to the field's max magnitude)
RTOL_P2G = 1e-4      # u, v, weights, density, dye vs JS-exact: the reference's binary32 running
                     # sums cancel in some faces (seen: 1.3e-5 in the hourglass)
RTOL_P2G_FP64 = 3e-7  # the same fields vs the FP64 oracle: only the final binary32 rounding
RTOL_RHSB = 1e-5     # bubble rhs: FP64 vs binary32 accumulation
RTOL_SOLVE = 2e-3    # q after 3 warm-started PCG iterations (FP32 V-cycle rounding)


@pytest.fixture(scope="module")
def tq():
    import tanque_hip
    tanque_hip.load()
    return tanque_hip


SCENARIOS = {
    "dam": (lambda orc: dam(orc), 20),
    "copo": (lambda orc: copo(orc), 30),
    "bolha": (lambda orc: tanque_com_bolha(orc), 6),
    "ampulheta": (lambda orc: ampulheta(orc), 20),
    "dam160": (lambda orc: dam(orc, nx=160, ny=90, h=0.8 / 90, seed=0), 12),
    # odd row length: the smoother tiles move cells one at a time instead of as row pairs
    "dam161": (lambda orc: dam(orc, nx=161, ny=90, h=0.8 / 90, seed=0), 12),
}


def advance(orc, osim, frames):
    for _ in range(frames):
        dt = substep_dt(orc, osim)
        sub = int(round((1 / 60) / dt))
        osim.avancar(1 / 60, sub)


@pytest.mark.parametrize("lds", [0, 99], ids=["mg-lds", "mg-multiblock"])
@pytest.mark.parametrize("name", list(SCENARIOS))
def test_phase_by_phase(oracle, tq, name, lds):
    """lds=0: the small grids' V-cycle runs in one workgroup; lds=99 keeps every level but the
    coarsest on the multi-workgroup kernels."""
    build, frames = SCENARIOS[name]
    osim = build(oracle)
    advance(oracle, osim, frames)
    gsim = hip_sim_like(tq, osim, mg_min_lds_level=lds)
    dt = substep_dt(oracle, osim)
    log("\n[%s] n=%d dt=%.6f" % (name, osim.n, dt))
    bad = []

    def run(ph_list, gph):
        clone_state(osim, gsim)
        for ph in ph_list:
            osim.fase(ph, dt)
        gsim.fase(gph, dt)

    # frame start: the air tracker (atualizarAr)
    log(" phase 12 tracker")
    run([12], 12)
    for f, nm in ((18, "rotulo"), (23, "volumeAlvo"), (24, "vivo"), (19, "pressaoBolha")):
        if cmp_exact(nm, gsim.read(f), osim.read(f)):
            bad.append(nm)
    # P2G + gravity/boundary + types
    log(" phase 0-2 P2G")
    osim_before = osim_snapshot(osim)
    run([0, 1, 2], 0)
    for f, nm in ((7, "contagem"), (9, "tipo")):
        if cmp_exact(nm, gsim.read(f), osim.read(f)):
            bad.append(nm)
    for f, nm in ((0, "u"), (1, "v"), (2, "uPeso"), (3, "vPeso"), (6, "densidade"), (8, "corCelula")):
        if not cmp_close(nm, gsim.read(f), osim.read(f), RTOL_P2G):
            bad.append(nm)
    # the same phase on the FP64 oracle from the same state: the GPU's error vs FP64
    o64 = oracle_fp64_clone(oracle, osim_before)
    for ph in (0, 1, 2):
        o64.fase(ph, dt)
    for f, nm in ((0, "u"), (1, "v"), (2, "uPeso"), (3, "vPeso"), (6, "densidade")):
        e_g = rel_err(gsim.read(f), o64.read(f, True))
        e_o = rel_err(osim.read(f), o64.read(f, True))
        log("  vs FP64 %-10s gpu %.2e  js-exact %.2e" % (nm, e_g, e_o))
        if e_g > RTOL_P2G_FP64:
            bad.append(nm + "(fp64)")
    log(" phase 3 labels")
    run([3], 3)
    if cmp_exact("rotulo", gsim.read(18), osim.read(18)):
        bad.append("rotulo(prop)")
    log(" phase 4 surface")
    run([4], 4)
    for f, nm in ((13, "fase"), (14, "normalX"), (15, "normalY"), (16, "curvatura")):
        if cmp_exact(nm, gsim.read(f), osim.read(f)):
            bad.append(nm)
    log(" phase 5 air pressure")
    run([5], 5)
    if cmp_exact("qAr", gsim.read(17), osim.read(17)):
        bad.append("qAr")
    log(" phase 6 divergence")
    run([6], 6)
    for f, nm in ((11, "div"), (12, "q")):
        if cmp_exact(nm, gsim.read(f), osim.read(f)):
            bad.append(nm)
    log(" phase 7 bubble rhs")
    run([7], 7)
    if cmp_exact("volumeBolha", gsim.read(20), osim.read(20)):
        bad.append("volumeBolha")
    if cmp_exact("yBolha", gsim.read(22), osim.read(22)):
        bad.append("yBolha")
    if not cmp_close("rhsBolha", gsim.read(21), osim.read(21), RTOL_RHSB):
        bad.append("rhsBolha")
    log(" phase 8 pressure")
    run([8], 8)
    st = osim.solver_stats()
    gs = gsim.solver_stats()
    log("  oracle norms %s | gpu norm0 %.4e norm %.4e iters %d bubbles %d" %
        (["%.4e" % x for x in st["normas"]], gs["norm0"], gs["norm"], gs["iterations"], gs["active_bubbles"]))
    if abs(gs["norm0"] - st["normas"][0]) > 1e-4 * max(st["normas"][0], 1e-30):
        bad.append("norm0")
    for f, nm in ((12, "q"), (22, "yBolha"), (19, "pressaoBolha")):
        if not cmp_close(nm, gsim.read(f), osim.read(f), RTOL_SOLVE):
            bad.append(nm + "(solve)")
    log(" phase 9 gradient")
    run([9], 9)
    for f, nm in ((0, "u"), (1, "v"), (4, "uValida"), (5, "vValida")):
        if cmp_exact(nm, gsim.read(f), osim.read(f)):
            bad.append(nm + "(grad)")
    log(" phase 10 extrapolation")
    run([10], 10)
    for f, nm in ((0, "u"), (1, "v"), (4, "uValida"), (5, "vValida")):
        if cmp_exact(nm, gsim.read(f), osim.read(f)):
            bad.append(nm + "(extrap)")
    log(" phase 11 G2P")
    run([11], 11)
    Pg, Po = gsim.particulas(), osim.particulas()
    log("  n gpu %d oracle %d; vMax gpu %r oracle %r" % (len(Pg), len(Po), gsim.vMax, osim.vmax))
    if len(Pg) != len(Po):
        bad.append("n(g2p)")
    else:
        if cmp_exact("particles", Pg, Po):
            bad.append("particles")
    if gsim.vMax != osim.vmax:
        bad.append("vMax")
    assert not bad, bad


def test_first_frame_tracker(oracle, tq):
    """The tracker's first snapshot (primeira): every enclosed air region >= 4 cells is a bubble."""
    osim = ampulheta(oracle)
    gsim = hip_sim_like(tq, osim)
    gsim.write_particles(osim.particulas())  # types are not known yet on either side
    osim.fase(12)
    gsim.fase(12)
    bad = 0
    for f, nm in ((9, "tipo"), (18, "rotulo"), (23, "volumeAlvo"), (24, "vivo")):
        bad += cmp_exact(nm, gsim.read(f), osim.read(f))
    assert bad == 0
    assert int((osim.read(18) >= 2).sum()) > 0  # the scene does have bubbles
