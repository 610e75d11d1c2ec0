"""Scene presets used by the reference's physics tests, restated from js/modelos.js.

Test infrastructure: returns strokes and water entries in metres for a W x H tank.
"""
MOLDURA = 2


def raio_de_parede(h):  # modelos.js:9-11
    return max(1.6 * h, 0.006)


def linha(pts, raio, fechar=False):  # modelos.js:13-21
    t = []
    n = len(pts) if fechar else len(pts) - 1
    for k in range(n):
        a, b = pts[k], pts[(k + 1) % len(pts)]
        t.append((a[0], a[1], b[0], b[1], raio, False))
    return t


def palco(W, H, h):  # modelos.js:24-34
    m = MOLDURA * h
    largura = min(W - 2 * m, 1.7 * H)
    return {"largura": largura, "chao": m, "teto": H - m, "altura": H - 2 * m,
            "X": lambda f: (W - largura) / 2 + f * largura,
            "Y": lambda f: m + f * (H - 2 * m)}


def copo(W, H, h):  # modelos.js:42-55
    p, r = palco(W, H, h), raio_de_parede(h)
    cx = W / 2
    meia = min(p["largura"] * 0.14, p["altura"] * 0.2)
    base = p["chao"] + r
    topo = p["Y"](0.46)
    return {"tracos": linha([[cx - meia * 1.18, topo], [cx - meia, base], [cx + meia, base], [cx + meia * 1.18, topo]], r),
            "agua": [{"x0": cx - meia * 0.75, "y0": p["Y"](0.66), "x1": cx + meia * 0.75, "y1": p["Y"](0.97)}]}


def ampulheta(W, H, h):  # modelos.js:79-107
    p, r = palco(W, H, h), raio_de_parede(h)
    cx = W / 2
    meia = min(p["largura"] * 0.2, p["altura"] * 0.3)
    pescoco = r + 1.8 * h
    topo, meio, base = p["Y"](0.94), p["Y"](0.5), p["Y"](0.06)
    contorno = [
        [cx - meia, topo], [cx + meia, topo], [cx + meia, topo - (topo - meio) * 0.25],
        [cx + pescoco, meio],
        [cx + meia, base + (meio - base) * 0.25], [cx + meia, base], [cx - meia, base],
        [cx - meia, base + (meio - base) * 0.25], [cx - pescoco, meio],
        [cx - meia, topo - (topo - meio) * 0.25],
    ]
    folga = r + h
    dentro = [
        [cx - meia + folga, topo - folga], [cx + meia - folga, topo - folga],
        [cx + meia - folga, topo - (topo - meio) * 0.25],
        [cx + pescoco * 0.3, meio + (topo - meio) * 0.12],
        [cx - pescoco * 0.3, meio + (topo - meio) * 0.12],
        [cx - meia + folga, topo - (topo - meio) * 0.25],
    ]
    return {"tracos": linha(contorno, r, True),
            "agua": [{"poligono": dentro, "limiteY": meio + (topo - meio) * 0.72}]}


def vasos(W, H, h):  # modelos.js:187-208
    p, r = palco(W, H, h), raio_de_parede(h)
    xa, xb = p["X"](0.18), p["X"](0.82)
    topo = p["Y"](0.82)
    vao = p["Y"](0.1)
    d1, d2 = xa + (xb - xa) * 0.3, xa + (xb - xa) * 0.52
    tracos = (linha([[xa, topo], [xa, p["chao"]]], r) + linha([[xb, topo], [xb, p["chao"]]], r) +
              linha([[d1, topo], [d1, vao]], r) + linha([[d2, topo], [d2 + (xb - d2) * 0.18, vao]], r))
    return {"tracos": tracos, "agua": [{"x0": xa, "y0": p["Y"](0.02), "x1": d1, "y1": p["Y"](0.62)}]}


MODELOS = {"copo": copo, "ampulheta": ampulheta, "vasos": vasos}


def dentro_do_poligono(x, y, pts):  # modelos.js:214-221
    dentro = False
    j = len(pts) - 1
    for i in range(len(pts)):
        xi, yi = pts[i]
        xj, yj = pts[j]
        if (yi > y) != (yj > y) and x < ((xj - xi) * (y - yi)) / (yj - yi) + xi:
            dentro = not dentro
        j = i
    return dentro


def semear_agua(orc, entrada, h, absorcao, aleatorio):  # modelos.js:227-242
    import numpy as np
    if "poligono" not in entrada:
        return orc.semear_retangulo(entrada["x0"], entrada["y0"], entrada["x1"], entrada["y1"], h, absorcao, aleatorio)
    xs = [q[0] for q in entrada["poligono"]]
    ys = [q[1] for q in entrada["poligono"]]
    topo = min(max(ys), entrada.get("limiteY", float("inf")))
    bruto = orc.semear_retangulo(min(xs), min(ys), max(xs), topo, h, absorcao, aleatorio).reshape(-1, 12)
    keep = [k for k in range(len(bruto)) if dentro_do_poligono(float(bruto[k, 0]), float(bruto[k, 1]), entrada["poligono"])]
    return np.ascontiguousarray(bruto[keep]).reshape(-1)


def montar_modelo(orc, nome, nx, ny, h, aleatorio, sim_cls=None, **opts):
    """montarModelo (testes/fisica.test.js:21-29)."""
    cena = orc.Cena(nx, ny, h)
    m = MODELOS[nome](cena.largura, cena.altura, h)
    for t in m["tracos"]:
        cena.tracar(*t)
    sim = (sim_cls or orc.Simulacao)(cena, **opts)
    for a in m["agua"]:
        sim.adicionar(semear_agua(orc, a, h, (0, 0, 0), aleatorio), False)
    return cena, sim, m
