"""The frame-level particle management on the device against the oracle, from identical state:
aplicarFerramenta (fisica-cpu.js:129-152: the sponge removes the particles inside its radius
keeping the others' order, the hand blends their velocity toward its own, the dye sets their
absorption) and the particle cap of MotorCPU.quadro (motor.js:100-104: each pending list is cut
to the room left under maxParticulas before adicionar).  Particle records, counts and the
following substep compared bit for bit."""
import numpy as np
import pytest

from gpu_common import clone_state, cmp_exact, dam, hip_sim_like, log, substep_dt

pytestmark = pytest.mark.gpu

CAMPOS = 12


@pytest.fixture(scope="module")
def tq():
    import tanque_hip
    tanque_hip.load()
    return tanque_hip


TOOLS = [
    {"tipo": "esponja", "x": 0.18, "y": 0.12, "raio": 0.07},
    {"tipo": "mao", "x": 0.15, "y": 0.20, "raio": 0.09, "vx": 1.7, "vy": -0.6},
    {"tipo": "corante", "x": 0.22, "y": 0.10, "raio": 0.06, "absorcao": (0.9, 0.25, 0.05)},
]


@pytest.mark.parametrize("k", range(len(TOOLS)), ids=[t["tipo"] for t in TOOLS])
def test_tool_matches_oracle(oracle, tq, k):
    osim = dam(oracle)
    for _ in range(3):
        dt = substep_dt(oracle, osim)
        osim.avancar(1 / 60, int(round((1 / 60) / dt)))
    g = hip_sim_like(tq, osim)
    clone_state(osim, g)
    f = TOOLS[k]
    n0 = osim.n
    osim.ferramenta(f["tipo"], f["x"], f["y"], f["raio"], f.get("vx", 0.0), f.get("vy", 0.0),
                    f.get("absorcao", (0, 0, 0)))
    g.aplicarFerramenta(f)
    log("\n  %s: %d -> %d particles (oracle %d)" % (f["tipo"], n0, g.n, osim.n))
    assert g.n == osim.n
    if f["tipo"] == "esponja":
        assert osim.n < n0  # the sponge sits in the water
    assert not cmp_exact("particles after the tool", g.particulas(), osim.particulas())
    # the substep after the tool: positions bit-exact, the P2G sums within their fixed-point tolerance
    dt = substep_dt(oracle, osim)
    osim.subpasso(dt)
    g.subpasso(dt)
    assert g.n == osim.n
    assert np.array_equal(g.read(7), osim.read(7))  # cell counts of the substep's P2G
    err = float(np.max(np.abs(g.particulas()[:, :2].astype(np.float64) - osim.particulas()[:, :2])))
    log("  after one substep: max position difference %.3e m" % err)
    assert err < 1e-4


def test_tools_in_sequence(oracle, tq):
    """All three tools one after another (the dye, then the hand moving the dyed water, then the
    sponge taking part of it), with a substep between them."""
    osim = dam(oracle)
    dt = substep_dt(oracle, osim)
    osim.avancar(1 / 60, int(round((1 / 60) / dt)))
    g = hip_sim_like(tq, osim)
    clone_state(osim, g)
    for f in (TOOLS[2], TOOLS[1], TOOLS[0]):
        osim.ferramenta(f["tipo"], f["x"], f["y"], f["raio"], f.get("vx", 0.0), f.get("vy", 0.0),
                        f.get("absorcao", (0, 0, 0)))
        g.aplicarFerramenta(f)
        assert not cmp_exact("particles after " + f["tipo"], g.particulas(), osim.particulas())
        clone_state(osim, g)  # the substep between tools starts from identical state again
        dt = substep_dt(oracle, osim)
        osim.subpasso(dt)
        g.subpasso(dt)
        clone_state(osim, g)


def test_particle_cap_matches_motor(oracle, tq):
    """Lists beyond maxParticulas = ceil(4.6 nx ny): the device keeps the first (room left)
    particles of each list, as MotorCPU.quadro cuts them before adicionar (checks off, so only
    the cap decides); further lists add nothing."""
    nx, ny, h = 40, 30, 0.02
    cena = oracle.Cena(nx, ny, h)
    osim = oracle.Simulacao(cena)
    g = hip_sim_like(tq, osim)
    teto = int(np.ceil(nx * ny * 4.6))
    sem = oracle.Semente(3)
    # the first list hugs the frame (adicionar's wall rule drops part of each cut list), the
    # others stay 4 cells inside it (every particle passes), so the cap is reached exactly
    lists = [oracle.semear_retangulo(2 * h, 2 * h, nx * h - 2 * h, ny * h - 2 * h, h, (0, 0, 0), sem)]
    lists += [oracle.semear_retangulo(4 * h, 4 * h, nx * h - 4 * h, ny * h - 4 * h, h, (0, 0, 0), sem) for _ in range(3)]
    sizes = []
    for lst in lists:
        lst = np.asarray(lst, np.float32).reshape(-1)
        n0 = osim.n
        cabe = max(0, teto - osim.n) * CAMPOS  # motor.js:101
        if cabe > 0:
            osim.adicionar(lst[:cabe] if lst.size > cabe else lst, False)
        acc = g.adicionar(lst, False)
        sizes.append((lst.size // CAMPOS, acc, osim.n, g.n))
        assert acc == osim.n - n0
        assert g.n == osim.n
    log("\n  cap %d: lists (size, accepted on the device, oracle n, device n) %s" % (teto, sizes))
    assert osim.n == teto == g.n
    assert sizes[-1][1] == 0
    assert not cmp_exact("particles at the cap", g.particulas(), osim.particulas())
    # stepping at the cap keeps the two identical in count
    dt = substep_dt(oracle, osim)
    osim.subpasso(dt)
    g.subpasso(dt)
    assert g.n == osim.n
