"""The reference's own test suite (testes/cena.test.js, testes/fisica.test.js), run on the oracle.

Every assertion and threshold is the reference's; passing them pins the oracle's physics to the
behaviour the reference's authors tested.  The same scenarios run on the HIP path in
test_gpu_reference_suite.py.
"""
import math

import numpy as np
import pytest

from modelos import MODELOS, montar_modelo, semear_agua

MOLDURA = 2
SDF_MAX = 10
PINCEL_MIN_CELULAS = 1.3
PPC = 4


# ------------------------------------------------------------------ testes/cena.test.js
def test_moldura_e_parede_e_meio_livre(oracle):  # cena.test.js:9-18
    h = 0.01
    c = oracle.Cena(40, 30, h)
    assert c.celula_solida(0, 0)
    assert c.celula_solida(MOLDURA - 1, 15)
    assert not c.celula_solida(MOLDURA, 15)
    assert not c.celula_solida(20, 15)
    assert c.distancia(0.2, 0.15) == SDF_MAX
    assert abs(c.distancia(MOLDURA * h + 3 * h, 0.15) - 3) < 0.05


def test_distancia_ate_traco_bate_com_geometria(oracle):  # cena.test.js:20-32
    h = 0.01
    c = oracle.Cena(60, 60, h)
    c.tracar(0.3, 0.3, 0.3, 0.3, 0.02)
    for d in (0.03, 0.045, 0.06, 0.08):
        esperado = (d - 0.02) / h
        assert abs(c.distancia(0.3 + d, 0.3) - esperado) < 0.35
    assert c.distancia(0.3, 0.3) < -1.5
    assert c.celula_solida(30, 30)
    nx, ny = c.normal(0.35, 0.3)
    assert nx > 0.99 and abs(ny) < 0.05


def test_apagar_nunca_apaga_moldura(oracle):  # cena.test.js:34-44
    h = 0.01
    c = oracle.Cena(40, 30, h)
    c.tracar(0.1, 0.1, 0.3, 0.1, 0.015)
    assert c.celula_solida(20, 10)
    c.tracar(0, 0.1, 0.4, 0.1, 0.03, True)
    c.tracar(0, 0, 0.4, 0.3, 0.05, True)
    assert not c.celula_solida(20, 10)
    assert c.celula_solida(0, 10) and c.celula_solida(39, 10) and c.celula_solida(20, 0)


def test_sdf_incremental_igual_ao_refeito(oracle):  # cena.test.js:46-66
    h = 0.01
    aleatorio = oracle.Semente(7)
    a = oracle.Cena(70, 50, h)
    for _ in range(25):
        t = (aleatorio() * 0.7, aleatorio() * 0.5, aleatorio() * 0.7, aleatorio() * 0.5,
             0.013 + aleatorio() * 0.03)
        apagar = aleatorio() < 0.25
        a.tracar(*t, apagar)
    b = oracle.Cena(70, 50, h)
    b.refazer(a.tracos)
    assert np.max(np.abs(a.sdf - b.sdf)) < 1e-5
    assert np.array_equal(a.solido, b.solido)


# ------------------------------------------------------------------ testes/fisica.test.js
def velocidade_maxima(P):  # apoio.js:18-22
    return float(np.max(np.hypot(P[:, 2].astype(np.float64), P[:, 3].astype(np.float64)))) if len(P) else 0.0


def menor_distancia_a_parede(sim):  # apoio.js:24-28
    P = sim.particulas()
    return min(sim.cena.distancia(float(x), float(y)) for x, y in P[:, :2])


@pytest.mark.parametrize("nome,tensao,limite", [("sem tensao", 0.0, 0.01), ("com tensao", 0.0728, 0.02)])
def test_agua_parada_fica_parada(oracle, nome, tensao, limite):  # fisica.test.js:31-43
    h = 0.01
    cena = oracle.Cena(60, 40, h)
    sim = oracle.Simulacao(cena, tensao=tensao)
    m = MOLDURA * h
    sim.adicionar(oracle.semear_retangulo(m, m, 0.6 - m, 0.2, h, (0, 0, 0), oracle.Semente(3)), False)
    oracle.rodar(sim, 3)
    assert velocidade_maxima(sim.particulas()) < limite


def test_tensao_arredonda_gota_quadrada(oracle):  # fisica.test.js:45-66
    h, n = 0.002, 70
    lado = 32 * h
    centro = (n * h) / 2

    def medir(tensao):
        cena = oracle.Cena(n, n, h)
        sim = oracle.Simulacao(cena, gravidade=(0.0, 0.0), tensao=tensao)
        sim.adicionar(oracle.semear_retangulo(centro - lado / 2, centro - lado / 2, centro + lado / 2,
                                              centro + lado / 2, h, (0, 0, 0), oracle.Semente(8)), False)
        oracle.rodar(sim, 1.2)
        P = sim.particulas().astype(np.float64)
        cx, cy = P[:, 0].mean(), P[:, 1].mean()
        return float(np.max(np.abs(P[:, 0] - cx) + np.abs(P[:, 1] - cy))) / lado

    com, sem = medir(0.0728), medir(0.0)
    assert sem > 0.93
    assert com < 0.88


def test_copo_nao_vaza_nem_perde_agua(oracle):  # fisica.test.js:68-76
    _, sim, _ = montar_modelo(oracle, "copo", 90, 60, 0.01, oracle.Semente(11))
    n0 = sim.n
    assert n0 > 1000
    pior = [math.inf]
    oracle.rodar(sim, 4, lambda t: pior.__setitem__(0, min(pior[0], menor_distancia_a_parede(sim))))
    assert sim.n == n0
    assert pior[0] > -0.05


def test_parede_fina_segura_jato(oracle):  # fisica.test.js:78-94
    h = 0.01
    cena = oracle.Cena(80, 60, h)
    cena.tracar(0.2, 0, 0.62, 0.6, PINCEL_MIN_CELULAS * h)
    sim = oracle.Simulacao(cena)
    jato = oracle.semear_retangulo(0.03, 0.35, 0.12, 0.55, h, (0, 0, 0), oracle.Semente(5)).reshape(-1, 12)
    jato[:, 2] = 3.5
    jato[:, 3] = -1
    sim.adicionar(jato.reshape(-1), False)
    vazou = [0]

    def conta(t):
        P = sim.particulas().astype(np.float64)
        vazou[0] += int(np.sum((P[:, 0] - 0.2) * 0.6 - P[:, 1] * 0.42 > 0))

    oracle.rodar(sim, 3, conta)
    assert vazou[0] == 0


def test_volume_conservado_no_rompimento(oracle):  # fisica.test.js:96-116
    h, nx, ny = 0.01, 80, 60
    cena = oracle.Cena(nx, ny, h)
    sim = oracle.Simulacao(cena)
    m = MOLDURA * h
    sim.adicionar(oracle.semear_retangulo(m, m, 0.3, 0.42, h, (0, 0, 0), oracle.Semente(9)), False)
    area = sim.area()
    oracle.rodar(sim, 10)
    P = sim.particulas()
    topo = np.zeros(nx, np.float32)
    for x, y in P[:, :2]:
        i = int(math.floor(float(x) / h))
        topo[i] = max(topo[i], y)
    nivel = float(np.mean(topo[MOLDURA + 3:nx - MOLDURA - 3].astype(np.float64))) - m
    esperado = area / (nx * h - 2 * m)
    assert abs(nivel - esperado) < 0.06 * esperado + 0.5 * h


def niveis_vasos(be, semente=21):
    """fisica.test.js:118-138: the three vessels' mean water levels over t >= 12 s (the highest
    particle between each pair of walls, 2.5 cells in from both)."""
    h = 0.01
    cena, sim, m = montar_modelo(be, "vasos", 100, 64, h, be.Semente(semente))
    tr = m["tracos"]
    xs = [tr[0][0], tr[2][0], (tr[3][0] + tr[3][2]) / 2, tr[1][0]]
    niveis = [[], [], []]

    def medir(t):
        if t < 12:
            return
        P = sim.particulas().astype(np.float64)
        for k in range(3):
            sel = (P[:, 0] > xs[k] + 2.5 * h) & (P[:, 0] < xs[k + 1] - 2.5 * h)
            niveis[k].append(float(P[sel, 1].max()) if sel.any() else 0.0)

    be.rodar(sim, 16, medir)
    return [sum(l) / len(l) for l in niveis]


@pytest.mark.slow
def test_vasos_comunicantes(oracle):  # fisica.test.js:118-138
    media = niveis_vasos(oracle)
    assert max(media) - min(media) < 2 * 0.01


def test_densidade_perto_do_repouso(oracle):  # fisica.test.js:140-159
    _, sim, _ = montar_modelo(oracle, "copo", 90, 60, 0.01, oracle.Semente(4))
    oracle.rodar(sim, 5)
    cont, sol = sim.contagem, sim.solido
    nx, ny = sim.nx, sim.ny
    soma = n = 0
    for j in range(2, ny - 2):
        for i in range(2, nx - 2):
            ok = True
            for b in (-1, 0, 1):
                for a in (-1, 0, 1):
                    c = i + a + (j + b) * nx
                    if cont[c] == 0 or sol[c]:
                        ok = False
            if ok:
                soma += int(cont[i + j * nx]); n += 1
    assert n > 50
    assert abs(soma / n - PPC) < 0.4


def tanque_com_bolha(oracle, ar):  # fisica.test.js:163-177
    h = 0.01
    cena = oracle.Cena(60, 60, h)
    sim = oracle.Simulacao(cena, ar=ar)
    m = MOLDURA * h
    agua = oracle.semear_retangulo(m, m, 0.6 - m, 0.45, h, (0, 0, 0), oracle.Semente(5)).reshape(-1, 12)
    x, y = agua[:, 0].astype(np.float64), agua[:, 1].astype(np.float64)
    fora = agua[~((x - 0.3) ** 2 + (y - 0.12) ** 2 < 0.05 ** 2)]
    sim.adicionar(fora.reshape(-1), False)
    return sim


def medir_bolhas(sim):  # fisica.test.js:179-188
    tipo, rot = sim.tipo, sim.rotulo
    sel = (tipo == 2) & (rot >= 2)
    n = int(sel.sum())
    if not n:
        return {"celulas": 0, "altura": None}
    j = np.nonzero(sel)[0] // sim.nx
    return {"celulas": n, "altura": float(np.sum((j + 0.5) * sim.h)) / n}


def test_bolha_sobe_mantendo_volume(oracle):  # fisica.test.js:190-200
    sim = tanque_com_bolha(oracle, True)
    oracle.rodar(sim, 0.1)
    inicio = medir_bolhas(sim)
    oracle.rodar(sim, 0.5)
    depois = medir_bolhas(sim)
    v0 = math.pi * 25
    assert inicio["celulas"] > 0.8 * v0
    assert depois["celulas"] > 0.75 * v0
    assert depois["altura"] - inicio["altura"] > 0.1


def test_remontar_nao_perde_bolha(oracle):  # fisica.test.js:202-222
    sim = tanque_com_bolha(oracle, True)
    m = MOLDURA * sim.h
    cheio = oracle.semear_retangulo(m, m, 0.6 - m, 0.45, sim.h, (0, 0, 0), oracle.Semente(6))
    sim.limpar_agua()
    sim.adicionar(cheio, False)
    oracle.rodar(sim, 0.1)
    com_bolha = tanque_com_bolha(oracle, True)
    nova = com_bolha.p.copy()
    sim.limpar_agua()
    sim.adicionar(nova, False)
    oracle.rodar(sim, 0.1)
    assert medir_bolhas(sim)["celulas"] > 0.8 * math.pi * 25


def test_sem_ar_buraco_fecha(oracle):  # fisica.test.js:224-230
    sim = tanque_com_bolha(oracle, False)
    oracle.rodar(sim, 0.3)
    tipo = sim.tipo.reshape(sim.ny, sim.nx)
    assert int(np.sum(tipo[3:25, 20:40] == 2)) < 10


def test_ampulheta_com_ar_esvazia_devagar(oracle):  # fisica.test.js:232-250
    def em_cima(ar):
        h = 0.01
        cena = oracle.Cena(60, 90, h)
        m = MODELOS["ampulheta"](cena.largura, cena.altura, h)
        for t in m["tracos"]:
            cena.tracar(*t)
        sim = oracle.Simulacao(cena, ar=ar)
        for a in m["agua"]:
            sim.adicionar(semear_agua(oracle, a, h, (0, 0, 0), oracle.Semente(2)), False)
        n0 = sim.n
        oracle.rodar(sim, 1)
        P = sim.particulas()
        return float(np.sum(P[:, 1].astype(np.float64) > cena.altura / 2)) / n0

    com, sem = em_cima(True), em_cima(False)
    assert sem < 0.1
    assert com > 0.4
