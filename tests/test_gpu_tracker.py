"""The device air tracker (union-find components, numbering, volume shares) against the oracle
on random air masks and on a split/merge sequence: labels and volume targets bit-exact."""
import numpy as np
import pytest
from scipy import ndimage

from gpu_common import log

pytestmark = pytest.mark.gpu

AGUA, AR, SOLIDO = 1, 2, 0
MOLDURA = 2


@pytest.fixture(scope="module")
def tq():
    import tanque_hip
    tanque_hip.load()
    return tanque_hip


def frame_types(nx, ny, air):
    t = np.where(air, AR, AGUA).astype(np.uint8)
    t[:MOLDURA, :] = SOLIDO
    t[-MOLDURA:, :] = SOLIDO
    t[:, :MOLDURA] = SOLIDO
    t[:, -MOLDURA:] = SOLIDO
    return t


def both(oracle, tq, nx, ny):
    ocena = oracle.Cena(nx, ny, 0.01)
    osim = oracle.Simulacao(ocena, mode=0, ar=True)
    osim.reiniciar_ar()
    gcena = tq.Cena(nx, ny, 0.01)
    gsim = tq.SimulacaoHIP(gcena, ar=True)
    gsim.reiniciarAr()
    return osim, gsim


def update(oracle, osim, gsim, tipo):
    osim.write(oracle.F_TIPO, tipo.reshape(-1))
    osim.set_tipos_prontos(True)
    osim.fase(oracle.PH_AR)
    gsim.write(9, tipo.reshape(-1))
    gsim.fase(12)
    return (osim.read(oracle.F_ROT), gsim.read(18).astype(np.uint32),
            osim.read(oracle.F_VALVO), gsim.read(23))


@pytest.mark.parametrize("seed", range(6))
def test_random_masks_bit_exact(oracle, tq, seed):
    nx, ny = 96, 72
    rng = np.random.default_rng(100 + seed)
    osim, gsim = both(oracle, tq, nx, ny)
    for frame in range(3):  # first frame, then two label histories
        field = ndimage.gaussian_filter(rng.standard_normal((ny, nx)), 1.0 + 0.5 * frame)
        ro, rg, vo, vg = update(oracle, osim, gsim, frame_types(nx, ny, field > 0.1))
        nb = len(np.unique(ro[ro >= 2]))
        log("\n  seed %d frame %d: %d bubbles, labels equal %s, volumes equal %s" % (
            seed, frame, nb, np.array_equal(ro, rg), np.array_equal(vo, vg)))
        assert np.array_equal(ro, rg)
        assert np.array_equal(vo, vg)


def test_split_and_merge_bit_exact(oracle, tq):
    nx, ny = 40, 30
    osim, gsim = both(oracle, tq, nx, ny)
    air = np.zeros((ny, nx), bool)
    air[8:14, 10:20] = True
    split = air.copy()
    split[8:14, 14] = False
    for a in (air, split, air, split):
        ro, rg, vo, vg = update(oracle, osim, gsim, frame_types(nx, ny, a))
        assert np.array_equal(ro, rg) and np.array_equal(vo, vg)
