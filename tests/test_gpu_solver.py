"""The pressure solve in the residual-terminated mode (PCG to a relative residual of 1e-6),
FP32 and mixed-precision multigrid, against the oracle run to the same tolerance, plus the
FP64 oracle for the error statements.

The stop test is on the PCG's recursive residual (an FP64 norm of the FP32 residual vector), the
quantity the reference's own solver reports (pressao-cpu.js:342-343, 386), followed by the true
residual: in this mode the device keeps the iterate as a compensated pair of binary32 words
(q + q_lo), recomputes ||b - A[q + q_lo; y]|| in binary64 at the stop and restarts from it while
it is above tol.  Each solve's iterate must reach tol in that true residual.  The binary32 q the
solve hands on keeps the storage floor of binary32 pressures (reported beside it).
"""
import numpy as np
import pytest

from gpu_common import (ampulheta, clone_state, dam, hip_sim_like, log, oracle_fp64_clone, osim_snapshot, rel_err,
                        substep_dt, tanque_com_bolha, tanque_com_bolhas)

pytestmark = pytest.mark.gpu

TOL = 1e-6
RTOL_Q = 1e-4          # q of two solves that both reached 1e-6 (FP32 storage)
RTOL_Q_MIXED = 2e-4    # the mixed-precision preconditioner takes a different path to 1e-6


@pytest.fixture(scope="module")
def tq():
    import tanque_hip
    tanque_hip.load()
    return tanque_hip


def prepared(orc, build, frames):
    osim = build(orc)
    for _ in range(frames):
        dt = substep_dt(orc, osim)
        osim.avancar(1 / 60, int(round((1 / 60) / dt)))
    osim.set_tol(TOL, 200)
    return osim


@pytest.mark.parametrize("name,build,frames", [
    ("dam", lambda o: dam(o), 15),
    ("bolha", lambda o: tanque_com_bolha(o), 6),
    ("ampulheta", lambda o: ampulheta(o), 20),
    ("dam160", lambda o: dam(o, nx=160, ny=90, h=0.8 / 90, seed=0), 12),
    # odd row length: the smoother tiles and the update move cells one at a time, not as row pairs
    ("dam161", lambda o: dam(o, nx=161, ny=90, h=0.8 / 90, seed=0), 12),
    # 256x144: level 1 (128x72) is large enough for FP16 storage in the mixed preconditioner
    ("dam256", lambda o: dam(o, nx=256, ny=144, h=1.0 / 144, seed=0), 8),
    # twelve bubbles: more than the per-block reduction slots, the rest in 128-bit fixed point
    ("bolhas12", lambda o: tanque_com_bolhas(o), 2),
    # 1024x640 (655k cells): 640 spmv and update blocks, their partials summed by every consumer block
    ("dam1024", lambda o: dam(o, nx=1024, ny=640, h=0.8 / 640, seed=0), 1),
    # 2048x1152 (2.36M cells): spmv and the update launch capped grids whose blocks walk several
    # chunks after one fold; the z partials (one per up-kernel block) are pre-reduced
    ("dam2048", lambda o: dam(o, nx=2048, ny=1152, h=0.8 / 1152, seed=0), 1),
])
def test_solve_to_tolerance(oracle, tq, name, build, frames):
    osim = prepared(oracle, build, frames)
    dt = substep_dt(oracle, osim)
    # bring the oracle to the solve: P2G .. bubble rhs
    for ph in (0, 1, 2, 3, 4, 5, 6, 7):
        osim.fase(ph, dt)
    snap = osim_snapshot(osim)
    o64 = oracle_fp64_clone(oracle, snap)
    o64.set_tol(TOL, 200)
    o64.fase(8, dt)
    s64 = o64.solver_stats()
    q64 = o64.read(12, True)
    res = {}
    for prec in (tq.PREC_FP32, tq.PREC_MIXED, tq.PREC_MIXED_B, tq.PREC_MIXED_X):
        g = hip_sim_like(tq, osim, iteracoes=200, tol=TOL, precision=prec)
        clone_state(osim, g)
        g.fase(8, dt)
        st = g.solver_stats(true_residual=True)
        res[prec] = (g.read(12), g.read(22), st)
    osim.fase(8, dt)
    so = osim.solver_stats()
    qo = osim.read(12)
    log("\n[%s] oracle: it %d rec %.2e true %.2e | fp64: it %d rec %.2e true %.2e" % (
        name, so["iters"], so["normas"][-1] / so["norm_b"], so["true_resid"] / so["norm_b"], s64["iters"],
        s64["normas"][-1] / s64["norm_b"], s64["true_resid"] / s64["norm_b"]))
    bad = []
    for prec, lab in ((tq.PREC_FP32, "fp32"), (tq.PREC_MIXED, "mixed"), (tq.PREC_MIXED_B, "mix-b"),
                      (tq.PREC_MIXED_X, "mix-x")):
        q, y, st = res[prec]
        rec = st["norm"] / st["norm_b"] if st["norm_b"] > 0 else 0
        tru = st["true_resid"] / st["norm_b"] if st["norm_b"] > 0 else 0
        tri = st["true_resid_iterate"] / st["norm_b"] if st["norm_b"] > 0 else 0
        e_or, e_64 = rel_err(q, qo), rel_err(q, q64)
        log("  gpu %-5s it %3d (first %3d, restarts %d) rec %.2e true %.2e (binary32 q %.2e) | q vs oracle %.2e "
            "vs fp64 %.2e | norm_b %.4e vs %.4e" % (lab, st["iterations"], st["first_iterations"], st["restarts"], rec,
                                                   tri, tru, e_or, e_64, st["norm_b"], so["norm_b"]))
        if not rec <= TOL:
            bad.append(lab + " not converged")
        if not (0 <= tri <= TOL):
            bad.append(lab + " true residual of the iterate %.2e" % tri)
        if abs(st["norm_b"] - so["norm_b"]) > 1e-5 * so["norm_b"]:
            bad.append(lab + " norm_b")
        if e_64 > (RTOL_Q if prec == tq.PREC_FP32 else RTOL_Q_MIXED):
            bad.append(lab + " q vs fp64 %.2e" % e_64)
        if prec == tq.PREC_FP32 and abs(st["first_iterations"] - so["iters"]) > 2:
            bad.append("fp32 iterations %d vs oracle %d" % (st["first_iterations"], so["iters"]))
    assert not bad, bad


def test_mixed_matches_fp32_over_steps(tq, oracle):
    """A few frames with the mixed preconditioner stay within tolerance of the FP32 run."""
    osim = dam(oracle)
    P0 = osim.particulas().copy()
    out = {}
    for prec in (tq.PREC_FP32, tq.PREC_MIXED, tq.PREC_MIXED_B, tq.PREC_MIXED_X):
        g = hip_sim_like(tq, osim, iteracoes=200, tol=TOL, precision=prec)
        g.write_particles(P0)
        for _ in range(10):
            g.avancar(1 / 60, 4)
        out[prec] = g.particulas()
    a, b = out[tq.PREC_FP32], out[tq.PREC_MIXED]
    assert len(a) == len(b)
    d = float(np.max(np.abs(a[:, :2].astype(np.float64) - b[:, :2])))
    log("\n  mixed vs fp32 after 10 frames: max position difference %.3e m (n=%d)" % (d, len(a)))
    assert d < 5e-3  # chaotic FLIP divergence over 40 substeps stays far below a cell (1 cm)


@pytest.mark.parametrize("prec", ["fp32", "mixed", "mix-b", "mix-x"])
@pytest.mark.parametrize("name,build", [
    ("bolha", lambda o: tanque_com_bolha(o)),
    ("dam160", lambda o: dam(o, nx=160, ny=90, h=0.8 / 90, seed=0)),
    ("dam1024", lambda o: dam(o, nx=1024, ny=640, h=0.8 / 640, seed=0)),
    ("dam2048", lambda o: dam(o, nx=2048, ny=1152, h=0.8 / 1152, seed=0)),  # capped grids, pre-reduced partials
])
def test_graph_replay_matches_eager(tq, oracle, prec, name, build):
    """PCG iterations replayed from the captured HIP graph give bit-identical states to eager
    launches, over many replays with the substep's other kernels and host syncs in between."""
    p = {"fp32": tq.PREC_FP32, "mixed": tq.PREC_MIXED, "mix-b": tq.PREC_MIXED_B, "mix-x": tq.PREC_MIXED_X}[prec]
    osim = build(oracle)
    P0 = osim.particulas().copy()
    out = {}
    for graph in (False, True):
        g = hip_sim_like(tq, osim, iteracoes=200, tol=TOL, precision=p, use_graph=graph)
        g.write_particles(P0)
        for _ in range(6):
            g.avancar(1 / 60, 3)
        out[graph] = (g.particulas(), g.read(12), g.read(22), g.solver_stats())
    a, b = out[False], out[True]
    log("\n  graph vs eager (%s %s): iterations %d / %d" % (name, prec, a[3]["total_iterations"],
                                                          b[3]["total_iterations"]))
    assert a[3]["total_iterations"] == b[3]["total_iterations"]
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("name,build,frames", [
    ("bolha", lambda o: tanque_com_bolha(o), 6),        # 60x60: level 1 (30x30) above a 15x15 and an 8x8 level
    ("dam160", lambda o: dam(o, nx=160, ny=90, h=0.8 / 90, seed=0), 12),
    ("dam256", lambda o: dam(o, nx=256, ny=144, h=1.0 / 144, seed=0), 8),
])
@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_dense_vcycle_matches_workgroup_vcycle(oracle, tq, name, build, frames, prec):
    """The V-cycle of the first one-workgroup level applied as this solve's dense matrix gives the
    PCG iterates of the same V-cycle run in one workgroup, to FP32 rounding: q after the
    reference's fixed iteration counts (3 and 10) and the iteration count to 1e-6."""
    p = {"fp32": tq.PREC_FP32, "mixed": tq.PREC_MIXED}[prec]
    osim = prepared(oracle, build, frames)
    dt = substep_dt(oracle, osim)
    for ph in range(8):
        osim.fase(ph, dt)
    bad = []
    for its, tol in ((3, 0.0), (10, 0.0), (200, TOL)):
        res = {}
        for dense in (True, False):
            g = hip_sim_like(tq, osim, iteracoes=its, tol=tol, precision=p, mg_dense=dense)
            clone_state(osim, g)
            g.timing_enable(True)
            g.fase(8, dt)
            names = set(g.timing_read())
            res[dense] = (g.read(12), g.solver_stats()["iterations"], names)
        (qd, itd, nd), (qw, itw, nw) = res[True], res[False]
        e = rel_err(qd, qw)
        log("\n  [%s %s] %3d its tol %g: dense vs workgroup q %.2e, iterations %d / %d" % (
            name, prec, its, tol, e, itd, itw))
        if "mg_coarse_mv" not in nd or "mg_vcycle_op" not in nd or "mg_coarse_wg" in nd:
            bad.append("dense path not taken: %s" % sorted(nd))
        if "mg_coarse_wg" not in nw or "mg_coarse_mv" in nw:
            bad.append("workgroup path not taken: %s" % sorted(nw))
        if e > 1e-4:
            bad.append("%d its: q differs by %.2e" % (its, e))
        if abs(itd - itw) > 1:
            bad.append("iterations %d vs %d" % (itd, itw))
    assert not bad, bad


def test_dense_operators_rebuilt_when_masks_change(tq, oracle):
    """The dense coarse operators are rebuilt only when the one-workgroup levels' masks change.
    Over a dam break (fixed 3 iterations, the reference's mode) a run that reuses them tracks a
    run with the workgroup V-cycle to FP32 rounding, and so does a run that switches the matrix
    off and on again (the switch forces a rebuild).  A stale operator would change the
    preconditioner, which with 3 iterations moves q far beyond rounding."""
    osim = dam(oracle, nx=160, ny=90, h=0.8 / 90, seed=0)
    P0 = osim.particulas().copy()
    sims = {}
    for name, dense in (("reuse", True), ("workgroup", False), ("toggled", True)):
        g = hip_sim_like(tq, osim, iteracoes=3, mg_dense=dense)
        g.write_particles(P0)
        sims[name] = g
    worst = 0.0
    for f in range(8):
        if f == 3:
            sims["toggled"].set_options(mg_dense=False)
        if f == 5:
            sims["toggled"].set_options(mg_dense=True)
        for g in sims.values():
            g.avancar(1 / 60, 3)
        qw = sims["workgroup"].read(12)
        for name in ("reuse", "toggled"):
            e = rel_err(sims[name].read(12), qw)
            worst = max(worst, e)
            log("\n  frame %d %-7s vs workgroup: q %.2e" % (f, name, e))
    st = sims["reuse"].solver_stats()
    log("  operator builds %d over %d solves" % (st["operator_builds"], st["solves"]))
    assert 1 <= st["operator_builds"] <= st["solves"]
    assert worst < 1e-3, worst


def test_launch_caps(tq, oracle):
    """tq_set_launch: a 2.36M-cell solve with the spmv / update grids capped lower than the
    defaults walks more cells per block; the same caps twice give identical bits, other caps
    the same pressure within the solve's tolerance.  Unknown names and negative caps are refused."""
    osim = prepared(oracle, lambda o: dam(o, nx=2048, ny=1152, h=0.8 / 1152, seed=0), 1)
    dt = substep_dt(oracle, osim)
    for ph in (0, 1, 2, 3, 4, 5, 6, 7):
        osim.fase(ph, dt)
    out = {}
    for caps in ((0, 0), (256, 640), (256, 640), (4096, 4096)):
        g = hip_sim_like(tq, osim, iteracoes=200, tol=TOL)
        g.set_launch("spmv_grid", caps[0])
        g.set_launch("update_grid", caps[1])
        assert (g.get_launch("spmv_grid"), g.get_launch("update_grid")) == caps
        clone_state(osim, g)
        g.fase(8, dt)
        st = g.solver_stats()
        q = g.read(12)
        if caps in out:
            assert np.array_equal(q, out[caps][0]), caps
        out[caps] = (q, st["iterations"])
        log("\n  caps %s: %d iterations" % (caps, st["iterations"]))
        if caps == (4096, 4096):
            with pytest.raises(tq.TqError):
                g.set_launch("no_such_kernel", 1)
            with pytest.raises(tq.TqError):
                g.set_launch("spmv_grid", -1)
    q0 = out[(0, 0)][0]
    for caps, (q, it) in out.items():
        assert rel_err(q, q0) <= RTOL_Q, (caps, rel_err(q, q0))


@pytest.mark.parametrize("name,build,frames", [
    ("ampulheta", lambda o: ampulheta(o), 20),
    ("dam1024", lambda o: dam(o, nx=1024, ny=640, h=0.8 / 640, seed=0), 1),
])
def test_true_residual_restart(oracle, tq, name, build, frames):
    """Far below binary32 resolution (tol 1e-10) the recursive residual of the binary32 CG vectors
    drifts from the true one, so the solve restarts from the binary64 true residual of the
    compensated iterate (residual replacement): the iterate's true residual reaches tol, which
    the recursive residual alone cannot promise, and q agrees with the FP64 oracle solved as far."""
    tol = 1e-10
    osim = prepared(oracle, build, frames)
    dt = substep_dt(oracle, osim)
    for ph in range(8):
        osim.fase(ph, dt)
    snap = osim_snapshot(osim)
    o64 = oracle_fp64_clone(oracle, snap)
    o64.set_tol(tol, 400)
    o64.fase(8, dt)
    q64 = o64.read(12, True)
    g = hip_sim_like(tq, osim, iteracoes=400, tol=tol)
    clone_state(osim, g)
    g.fase(8, dt)
    st = g.solver_stats(true_residual=True)
    tri = st["true_resid_iterate"] / st["norm_b"]
    e64 = rel_err(g.read(12), q64)
    log("\n  [%s] tol %g: %d iterations (first segment %d), %d restarts, true residual of the iterate %.2e, "
        "binary32 q %.2e, q vs fp64 %.2e" % (name, tol, st["iterations"], st["first_iterations"], st["restarts"], tri,
                                            st["true_resid"] / st["norm_b"], e64))
    assert st["restarts"] >= 1
    assert 0 <= tri <= tol
    assert e64 < 1e-6
