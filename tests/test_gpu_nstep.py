"""After N substeps from one state: the HIP path against the JS-exact oracle, and both against
the FP64 oracle, with the pressure solved to a relative residual of 1e-6 on all three.

From a mid-run oracle state cloned into the HIP simulation and into an FP64-storage oracle, the
three run the same substeps; after 1, 4 and 16 of them the particle velocities and positions,
the pressure field and the bubble pressures are compared.  Each error is stated against the
FP64 oracle, for the HIP path and for the JS-exact oracle side by side: the HIP path's error
must stay within a small multiple of the JS-exact oracle's own (both store binary32; the HIP
path sums P2G in fixed point and solves in FP32 with FP64 dots).
"""
import numpy as np
import pytest

from gpu_common import (ampulheta, clone_state, dam, hip_sim_like, log, oracle_fp64_clone, osim_snapshot,
                        substep_dt, tanque_com_bolha)

pytestmark = pytest.mark.gpu

TOL = 1e-6
N_CHECK = (1, 4, 16)
# HIP error vs FP64 must not exceed this multiple of the JS-exact oracle's error vs FP64, plus a
# floor for fields where the JS-exact error happens to be ~0
FACTOR, FLOOR = 10.0, {"vel": 2e-5, "pos": 2e-6, "q": 2e-4, "yb": 2e-4}

SCENES = {
    "bolha": (lambda orc: tanque_com_bolha(orc), 6),
    "ampulheta": (lambda orc: ampulheta(orc), 20),
    "dam": (lambda orc: dam(orc), 20),
}


def advance(osim, orc, frames):
    for _ in range(frames):
        dt = substep_dt(orc, osim)
        osim.avancar(1 / 60, int(round((1 / 60) / dt)))


def rel(a, b):
    d = np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64))
    n = np.linalg.norm(np.asarray(b, np.float64))
    return d / n if n > 0 else d


@pytest.fixture(scope="module")
def tq():
    import tanque_hip
    tanque_hip.load()
    return tanque_hip


@pytest.mark.parametrize("name", list(SCENES))
def test_n_substeps_against_fp64(oracle, tq, name):
    orc = oracle
    build, frames = SCENES[name]
    osim = build(orc)
    advance(osim, orc, frames)
    snap = osim_snapshot(osim)
    o64 = oracle_fp64_clone(orc, snap, tol=TOL, iteracoes=200)
    osim.set_tol(TOL, 200)
    gsim = hip_sim_like(tq, osim, iteracoes=200, tol=TOL)
    clone_state(osim, gsim)
    dt = substep_dt(orc, osim)
    log("n-step %s: %d particles, dt %.5f s" % (name, osim.n, dt))
    h = osim.h
    for n in range(1, max(N_CHECK) + 1):
        osim.subpasso(dt)
        o64.subpasso(dt)
        gsim.subpasso(dt)
        if n not in N_CHECK:
            continue
        Pj, P6, Pg = osim.particulas(), o64.particulas(), gsim.particulas()
        assert len(Pj) == len(P6) == len(Pg), (name, n)
        tipo_j, tipo_g = osim.read(orc.F_TIPO), gsim.read(orc.F_TIPO)
        water = (tipo_j == orc.AGUA) & (tipo_g == orc.AGUA)
        err = {}
        for who, P, sim in (("js", Pj, osim), ("hip", Pg, gsim)):
            err[who] = {
                "vel": rel(P[:, 2:4], P6[:, 2:4]),
                "pos": float(np.abs(P[:, :2].astype(np.float64) - P6[:, :2]).max()) / h,
                "q": rel(sim.read(orc.F_Q)[water], o64.read(orc.F_Q)[water]),
                "yb": rel(sim.read(orc.F_YBOLHA), o64.read(orc.F_YBOLHA)),
            }
        log("  after %2d: vs FP64  velocity js %.2e hip %.2e | position (cells) js %.2e hip %.2e | q js %.2e hip %.2e"
            " | bubble y js %.2e hip %.2e | types differ in %d cells"
            % (n, err["js"]["vel"], err["hip"]["vel"], err["js"]["pos"], err["hip"]["pos"], err["js"]["q"],
               err["hip"]["q"], err["js"]["yb"], err["hip"]["yb"], int((tipo_j != tipo_g).sum())))
        for k in ("vel", "pos", "q", "yb"):
            assert err["hip"][k] <= FACTOR * err["js"][k] + FLOOR[k], (name, n, k, err)
