"""tq_autotune: the solver's launch shapes tuned on a tank's pressure system leave the tank's
state untouched, are applied and cached for the grid (a new tank of the same grid starts with
them), and solves with them reach the tolerance and the same pressure within it."""
import numpy as np
import pytest

from gpu_common import dam, hip_sim_like, log, rel_err

pytestmark = pytest.mark.gpu

KNOBS = ("spmv_grid", "update_grid", "split_cells", "pass_fold_grid")


@pytest.fixture(scope="module")
def tq():
    import tanque_hip
    tanque_hip.load()
    return tanque_hip


def test_autotune_keeps_state_and_caches(oracle, tq):
    osim = dam(oracle, nx=384, ny=216, h=0.8 / 216, seed=0)
    P0 = osim.particulas().copy()
    g = hip_sim_like(tq, osim, iteracoes=200, tol=1e-6)
    g.write_particles(P0)
    for _ in range(3):
        g.avancar(1 / 60, 2)
    q0, p0 = g.read(12).copy(), g.particulas().copy()
    st0 = g.solver_stats()
    r = g.autotune(10)
    log("\n  autotune 384x216: %s" % r)
    assert r["trials"] >= 2 and r["default_us"] > 0 and 0 < r["tuned_us"] <= r["default_us"]
    assert np.array_equal(g.read(12), q0) and np.array_equal(g.particulas(), p0)
    st1 = g.solver_stats()
    assert (st1["solves"], st1["total_iterations"]) == (st0["solves"], st0["total_iterations"])
    for k in KNOBS:
        assert g.get_launch(k) == r[k]
    # a new tank of the same grid starts with the cached choice
    g2 = hip_sim_like(tq, osim, iteracoes=200, tol=1e-6)
    for k in KNOBS:
        assert g2.get_launch(k) == r[k]
    g2.close()
    # stepping on: the tuned shapes solve to the tolerance, the pressure within it of the defaults'
    ref = hip_sim_like(tq, osim, iteracoes=200, tol=1e-6)
    for k in KNOBS:
        ref.set_launch(k, 0)
    ref.set_options(mg_dense=True)
    ref.write_particles(P0)
    for _ in range(3):
        ref.avancar(1 / 60, 2)
    assert np.array_equal(ref.read(12), q0)  # the same start as the tuned tank
    g.avancar(1 / 60, 2)
    ref.avancar(1 / 60, 2)
    st = g.solver_stats(true_residual=True)
    e = rel_err(g.read(12), ref.read(12))
    log("  after tuning: true residual %.2e, q vs default shapes %.2e" % (st["true_resid_iterate"] / st["norm_b"], e))
    assert st["true_resid_iterate"] <= 1e-6 * st["norm_b"]
    assert e < 1e-3
    g.close()
    ref.close()
