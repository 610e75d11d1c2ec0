"""Writes tests/golden/reference_runs.json: golden vectors of the reference's own CPU engine.

Run where the reference's sources are checked out (they are not part of this repository; the
committed fixture is what the tests read):
    python tests/golden/gen_reference_runs.py [/root/reference]

The reference (js/constantes.js, cena.js, fisica-cpu.js, pressao-cpu.js, ar.js, modelos.js,
testes/apoio.js) needs a newer JavaScript engine than the image's Node v12.  This script lowers
the few newer constructs it uses, in memory, without touching the arithmetic:
  * private names `#m` -> `_priv_m`
  * `X ??= Y` -> `(X != null ? X : (X = Y))`, `X ?? Y` -> `(X != null ? X : Y)` (X side-effect free)
  * `X?.y` -> `(X == null ? undefined : X.y)`
  * `Array.prototype.at` polyfilled
  * ES modules wrapped as functions returning their exports
and pipes the bundle, followed by the driver below, to `node -` on stdin: no copy of the
reference is written anywhere.  Each rewrite is counted and checked, so a changed reference
stops the script instead of being lowered wrongly.

The driver runs seeded scenes of the reference's own models (MODELOS with the LCG of
testes/apoio.js in place of Math.random for the maze and the water, semearAgua) with the reference tests' frame loop (fisica.test.js:12-19: frames of 1/60 s
split by passoMaximo), phase by phase in subpasso's order, and records after every phase a
SHA-1 of all the state the oracle exposes (the grid arrays, labels, bubble arrays, the tracker's
volumes and live flags, and the particles except the foam column), the foam column's own hash,
and the particles at a few checkpoints.  Only these vectors are committed.
"""
import base64
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "reference_runs.json")

MODULES = [  # (name, path, expected rewrites: ??=, ??, ?.)
    ("constantes", "js/constantes.js", (0, 1, 0)),
    ("cena", "js/cena.js", (0, 0, 0)),
    ("pressao", "js/pressao-cpu.js", (2, 3, 1)),
    ("ar", "js/ar.js", (2, 2, 0)),
    ("fisica", "js/fisica-cpu.js", (2, 5, 0)),
    ("modelos", "js/modelos.js", (0, 1, 0)),
    ("apoio", "testes/apoio.js", (0, 0, 0)),
]
PATH_TO_MOD = {"./constantes.js": "constantes", "./cena.js": "cena", "./pressao-cpu.js": "pressao", "./ar.js": "ar",
               "./fisica-cpu.js": "fisica", "./modelos.js": "modelos", "../js/fisica-cpu.js": "fisica"}

OPERAND = r"[A-Za-z_$][\w$]*(?:\.[A-Za-z_$][\w$]*|\[[^\[\]]*\]|\((?:[^()]|\([^()]*\))*\))*"


def lower(src, expect):
    n_assign = n_coal = n_opt = 0

    def sub_assign(m):
        nonlocal n_assign
        n_assign += 1
        x, y = m.group(1), m.group(2)
        return "(%s != null ? %s : (%s = %s))" % (x, x, x, y)

    # X ??= Y up to the end of the statement (one per line here)
    src = re.sub(r"(%s)\s*\?\?=\s*([^;\n]*);" % OPERAND, lambda m: sub_assign(m) + ";", src)

    def sub_coal(m):
        nonlocal n_coal
        n_coal += 1
        x, y = m.group(1), m.group(2)
        return "(%s != null ? %s : %s)" % (x, x, y)

    src = re.sub(r"(%s)\s*\?\?\s*(\[[^\[\]]*\]|[A-Za-z_$0-9][\w$.]*|-?\d+(?:\.\d+)?)" % OPERAND, sub_coal, src)

    def sub_opt(m):
        nonlocal n_opt
        n_opt += 1
        return "(%s == null ? undefined : %s.%s)" % (m.group(1), m.group(1), m.group(2))

    src = re.sub(r"([A-Za-z_$][\w$]*)\?\.([A-Za-z_$][\w$]*)", sub_opt, src)
    # private names (not inside quoted colour strings like '#d8eef0')
    src = re.sub(r"(?<!['\"\w])#([A-Za-z_][\w]*)", r"_priv_\1", src)
    got = (n_assign, n_coal, n_opt)
    if got != expect:
        raise SystemExit("lowering: expected %s rewrites, made %s" % (expect, got))
    if re.search(r"\?\?|\?\.[A-Za-z_]", src):
        raise SystemExit("lowering: a construct is left")
    return src


def as_function(name, src):
    """An ES module as a function of the modules loaded before it, returning its exports."""
    imports = []

    def sub_import(m):
        names = " ".join(m.group(1).split())
        mod = PATH_TO_MOD[m.group(2)]
        imports.append(mod)
        return "const {%s} = __m.%s;" % (names, mod)

    src = re.sub(r"import\s*\{([^}]*)\}\s*from\s*'([^']+)';", sub_import, src)
    if re.search(r"^\s*import\b", src, re.M):
        raise SystemExit("lowering: an import form is not handled in " + name)
    exported = re.findall(r"^export\s+(?:const|let|function\*?|class)\s+([A-Za-z_$][\w$]*)", src, re.M)
    # `export const A = 0, B = 1, C = 2;`
    for m in re.finditer(r"^export\s+const\s+([^;{}()\[\]\n]*);", src, re.M):
        for part in m.group(1).split(","):
            nm = part.split("=")[0].strip()
            if re.fullmatch(r"[A-Za-z_$][\w$]*", nm) and nm not in exported:
                exported.append(nm)
    src = re.sub(r"^export\s+", "", src, flags=re.M)
    return "__m.%s = (function () {\n%s\nreturn { %s };\n})();\n" % (name, src, ", ".join(exported))


PRELUDE = r"""'use strict';
const __m = {};
if (!Array.prototype.at) {
  Object.defineProperty(Array.prototype, 'at', { value: function (i) { i = Math.trunc(i) || 0; if (i < 0) i += this.length; return this[i]; } });
}
"""

DRIVER = r"""
const crypto = require('crypto');
const { Cena } = __m.cena;
const { SimulacaoCPU, CAMPOS, semearRetangulo } = __m.fisica;
const { MODELOS, semearAgua } = __m.modelos;
const { semente } = __m.apoio;
const { passoMaximo } = __m.constantes;

const QUADRO = 1 / 60;
function canon(a) { return Buffer.from(Float64Array.from(a).buffer); }
function sha(bufs) { const h = crypto.createHash('sha1'); for (const b of bufs) h.update(b); return h.digest('hex').slice(0, 16); }
function grid(sim) {
  const r = sim.rastreador;
  return [sim.u, sim.v, sim.uPeso, sim.vPeso, sim.uValida, sim.vValida, sim.densidade, sim.contagemCelula, sim.corCelula,
          sim.tipo, sim.solidoCelula, sim.div, sim.q, sim.fase, sim.normalX, sim.normalY, sim.curvatura, sim.qAr,
          sim.rotulo, sim.pressaoBolha, sim.volumeBolha, sim.rhsBolha, sim.yBolha, r.volumeAlvo, r.vivo];
}
function parts(sim, foam) {
  const n = sim.n, out = new Float64Array(foam ? n : n * 11);
  for (let k = 0; k < n; k++) {
    if (foam) { out[k] = sim.p[k * CAMPOS + 8]; continue; }
    for (let c = 0, q = 0; c < CAMPOS; c++) if (c !== 8) out[k * 11 + q++] = sim.p[k * CAMPOS + c];
  }
  return Buffer.from(out.buffer);
}
function record(sim) { return { s: sha(grid(sim).map(canon).concat([parts(sim, false)])), f: sha([parts(sim, true)]), n: sim.n }; }

function run(sc) {
  const cena = new Cena(sc.nx, sc.ny, sc.h);
  const rnd = semente(sc.semente);
  const m = MODELOS[sc.modelo].montar(cena.largura, cena.altura, sc.h, rnd);
  for (const t of m.tracos) cena.tracar(t.x0, t.y0, t.x1, t.y1, t.raio);
  cena.atualizar();
  const sim = new SimulacaoCPU(cena, sc.opcoes);
  for (const a of m.agua) sim.adicionar(semearAgua(a, sc.h, [0, 0, 0], rnd), false);
  const out = { sdf: Buffer.from(cena.sdf.buffer, cena.sdf.byteOffset, cena.sdf.byteLength).toString('base64'),
                p0: Buffer.from(sim.p.buffer, 0, sim.n * CAMPOS * 4).toString('base64'), n0: sim.n,
                frames: [], checkpoints: {} };
  let sub_index = 0;
  for (let f = 0; f < sc.quadros; f++) {
    // tools and added water between frames (fisica-cpu.js:92-111, 129-152), as the app applies them
    const ev = [];
    for (const e of (sc.eventos || []).filter((e) => e.quadro === f)) {
      let aceitas = null;
      if (e.tipo === 'agua') {
        const r = e.retangulo;
        aceitas = sim.adicionar(semearRetangulo(r[0], r[1], r[2], r[3], sc.h, e.absorcao, semente(e.semente)), e.checar);
      } else {
        sim.aplicarFerramenta(e.ferramenta);
      }
      ev.push(Object.assign({ aceitas }, record(sim)));
    }
    const sub = Math.ceil(QUADRO / passoMaximo(sim.h, sim.cena.altura, sim.vMax == null ? null : sim.vMax, QUADRO));
    const fr = { sub, ar: null, phases: [], eventos: ev };
    // avancar (fisica-cpu.js:155-159), phase by phase as subpasso (176-189) calls them
    if (sim.arAtivo) { sim.atualizarAr(); fr.ar = record(sim); }
    const d = QUADRO / sub;
    for (let s = 0; s < sub; s++) {
      const ph = [];
      sim._priv_paraGrade(); ph.push(record(sim));
      sim._priv_gravidadeEFronteira(d); ph.push(record(sim));
      sim._priv_tipos(); ph.push(record(sim));
      if (sim.arAtivo) sim._priv_propagarRotulos(); ph.push(record(sim));
      sim._priv_superficie(); ph.push(record(sim));
      sim._priv_pressaoDoAr(d); ph.push(record(sim));
      sim._priv_classificar(d); ph.push(record(sim));
      if (sim.arAtivo) sim._priv_ladoDireitoDasBolhas(d); ph.push(record(sim));
      sim._priv_pressao(d); ph.push(record(sim));
      sim._priv_gradiente(); ph.push(record(sim));
      sim._priv_extrapolar(); ph.push(record(sim));
      sim._priv_paraParticulas(d); ph.push(record(sim));
      fr.phases.push(ph);
      sub_index++;
      if (sc.checkpoints.includes(sub_index))
        out.checkpoints[sub_index] = Buffer.from(sim.p.buffer, 0, sim.n * CAMPOS * 4).toString('base64');
    }
    out.frames.push(fr);
  }
  return out;
}

const result = {};
for (const sc of SCENES) result[sc.nome] = Object.assign({ config: sc }, run(sc));
process.stdout.write(JSON.stringify(result));
"""

# seeded scenes: the reference's own models at small sizes (tank 1.42 x 0.80 m); checkpoints are
# substep indices after which the particles are stored whole
SCENES = [
    {"nome": "copo", "modelo": "copo", "nx": 64, "ny": 36, "h": 0.8 / 36, "semente": 1, "quadros": 45,
     "opcoes": {}, "checkpoints": [10, 40, 80]},
    {"nome": "ampulheta", "modelo": "ampulheta", "nx": 64, "ny": 36, "h": 0.8 / 36, "semente": 2, "quadros": 45,
     "opcoes": {}, "checkpoints": [10, 40, 80]},
    {"nome": "vasos", "modelo": "vasos", "nx": 64, "ny": 36, "h": 0.8 / 36, "semente": 3, "quadros": 30,
     "opcoes": {"tensao": 0}, "checkpoints": [20, 60]},
    {"nome": "funil", "modelo": "funil", "nx": 48, "ny": 27, "h": 0.8 / 27, "semente": 4, "quadros": 40,
     "opcoes": {"iteracoes": 6}, "checkpoints": [30]},
    {"nome": "cascata", "modelo": "cascata", "nx": 48, "ny": 27, "h": 0.8 / 27, "semente": 5, "quadros": 40,
     "opcoes": {}, "checkpoints": [30]},
    {"nome": "labirinto", "modelo": "labirinto", "nx": 64, "ny": 36, "h": 0.8 / 36, "semente": 6, "quadros": 30,
     "opcoes": {}, "checkpoints": [40]},
    # the tools and checked additions (A18): dye, hand, sponge, water poured onto water
    {"nome": "ferramentas", "modelo": "copo", "nx": 64, "ny": 36, "h": 0.8 / 36, "semente": 7, "quadros": 36,
     "opcoes": {}, "checkpoints": [60], "eventos": [
         {"quadro": 4, "tipo": "ferramenta", "ferramenta": {"tipo": "corante", "x": 0.71, "y": 0.3, "raio": 0.09,
                                                             "absorcao": [60, 26, 3]}},
         {"quadro": 8, "tipo": "ferramenta", "ferramenta": {"tipo": "mao", "x": 0.66, "y": 0.32, "raio": 0.08,
                                                             "vx": 1.4, "vy": 0.6}},
         {"quadro": 12, "tipo": "agua", "retangulo": [0.6, 0.25, 0.82, 0.45], "semente": 11, "checar": True,
          "absorcao": [2, 55, 42]},
         {"quadro": 16, "tipo": "ferramenta", "ferramenta": {"tipo": "esponja", "x": 0.75, "y": 0.28, "raio": 0.06}},
         {"quadro": 20, "tipo": "agua", "retangulo": [0.3, 0.5, 0.5, 0.7], "semente": 12, "checar": True,
          "absorcao": [1, 7, 60]},
         {"quadro": 24, "tipo": "ferramenta", "ferramenta": {"tipo": "mao", "x": 0.72, "y": 0.2, "raio": 0.1,
                                                             "vx": -2.0, "vy": 1.0}},
         {"quadro": 28, "tipo": "ferramenta", "ferramenta": {"tipo": "nenhuma"}},
     ]},
]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    bundle = [PRELUDE]
    for name, path, expect in MODULES:
        with open(os.path.join(ref, path), encoding="utf-8") as f:
            bundle.append(as_function(name, lower(f.read(), expect)))
    bundle.append("const SCENES = %s;\n" % json.dumps(SCENES))
    bundle.append(DRIVER)
    r = subprocess.run(["node", "-"], input="".join(bundle), capture_output=True, text=True, timeout=1800)
    if r.returncode != 0:
        sys.stderr.write(r.stderr[-4000:])
        raise SystemExit("node failed")
    data = json.loads(r.stdout)
    data["_about"] = ("Golden vectors of the reference's CPU engine (tests/golden/gen_reference_runs.py): per scene "
                      "the SDF and seeded particles, per frame the substep count and, after atualizarAr and after each "
                      "phase of each substep, SHA-1 (16 hex) of the state in float64 (s: grid arrays, labels, bubble "
                      "arrays, tracker volumes/live flags, particles without foam; f: foam), and particles at "
                      "checkpoint substeps (little-endian float32, base64).")
    with open(OUT, "w") as f:
        json.dump(data, f, separators=(",", ":"))
    print("wrote %s (%d bytes)" % (OUT, os.path.getsize(OUT)))


if __name__ == "__main__":
    main()
