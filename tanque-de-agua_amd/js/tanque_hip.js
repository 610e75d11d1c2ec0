'use strict';
// tanque_hip.js -- the MI355X substep for the reference's JavaScript host (CommonJS, Node >= 12).
//
// SimulacaoHIP keeps SimulacaoCPU's constructor and members (js/fisica-cpu.js:28-194): the same
// options, `n` (settable to truncate), `p` / `particulas()`, `vMax`, `gravidade`, the typed state
// arrays as read-only copies, and the methods atualizarParedes, adicionar, aplicarFerramenta,
// avancar, subpasso, atualizarAr, reiniciarAr, area.  It accepts the reference's own Cena
// (js/cena.js: nx, ny, h, sdf) or the Cena below.  MotorHIP has MotorCPU's surface
// (js/motor.js:79-120) without the 2D renderer.
const path = require('path');

const nativo = require(path.join(__dirname, '..', 'lib', 'tanque_node.node'));

const CAMPOS = 12;                      // floats per particle (fisica-cpu.js:26)
const PARTICULAS_MAX_POR_CELULA = 4.6;  // constantes.js
const CAMPO = {
  u: 0, v: 1, uPeso: 2, vPeso: 3, uValida: 4, vValida: 5, densidade: 6, contagemCelula: 7,
  corCelula: 8, tipo: 9, solidoCelula: 10, div: 11, q: 12, fase: 13, normalX: 14, normalY: 15,
  curvatura: 16, qAr: 17, rotulo: 18, pressaoBolha: 19, volumeBolha: 20, rhsBolha: 21, yBolha: 22,
  volumeAlvo: 23, vivo: 24,
};

/** Walls as an SDF on the reference's 2x2-per-cell lattice (js/cena.js), built by the library. */
class Cena {
  constructor(nx, ny, h) {
    this.nx = nx;
    this.ny = ny;
    this.h = h;
    this.largura = nx * h;
    this.altura = ny * h;
    this.sdf = nativo.cenaMoldura(nx, ny);
    this.tracos = [];
    this.linhasAlteradas = [0, 2 * ny - 1];
  }

  /** tracar (cena.js:43-72): a capsule stroke, `apagar` erases. */
  tracar(x0, y0, x1, y1, raio, apagar = false) {
    nativo.tracar(this.sdf, this.nx, this.ny, this.h, x0, y0, x1, y1, raio, apagar);
    this.tracos.push([x0, y0, x1, y1, raio, apagar]);
    this.linhasAlteradas = [0, 2 * this.ny - 1];
  }

  /** pegarLinhasAlteradas (cena.js:95-99). */
  pegarLinhasAlteradas() {
    const l = this.linhasAlteradas;
    this.linhasAlteradas = null;
    return l;
  }
}

/** semente (testes/apoio.js:4-9): the seeded LCG; the state is advanced by semearRetangulo. */
class Semente {
  constructor(s = 1) { this.estado = s >>> 0; }
  /** The next number in [0, 1) of the tests' LCG (testes/apoio.js:4-9): s = s*1664525 + 1013904223 mod 2^32. */
  proximo() {
    this.estado = (Math.imul(this.estado, 1664525) + 1013904223) >>> 0;
    return this.estado / 4294967296;
  }
}

/** semearRetangulo (fisica-cpu.js:705-727) with a Semente. */
function semearRetangulo(x0, y0, x1, y1, h, absorcao = [0, 0, 0], semente = new Semente(1)) {
  const r = nativo.semearRetangulo(x0, y0, x1, y1, h, absorcao, semente.estado);
  semente.estado = r.semente >>> 0;
  return r.particulas;
}

/** passoMaximo (constantes.js:167-175). */
function passoMaximo(h, altura, vMax, dtQuadro) {
  return nativo.passoMaximo(h, altura, vMax === undefined ? null : vMax, dtQuadro);
}

class SimulacaoHIP {
  constructor(cena, opcoes = {}) {
    this.cena = cena;
    this.nx = cena.nx;
    this.ny = cena.ny;
    this.h = cena.h;
    this._opcoes = Object.assign({}, opcoes);
    this._g = opcoes.gravidade !== undefined && opcoes.gravidade !== null ? opcoes.gravidade.slice() : [0, -9.81];
    this._n = new nativo.SimulacaoHIP(cena.nx, cena.ny, cena.h, opcoes);
    this.atualizarParedes();
  }

  /** atualizarParedes (fisica-cpu.js:80-85): every solid cell from the scene's SDF. */
  atualizarParedes(linha0 = 0, linha1 = 2 * this.ny - 1) { this._n.definirSdf(this.cena.sdf, linha0, linha1); }
  /** adicionar (fisica-cpu.js:92-111) -> accepted count. */
  adicionar(lista, checarCheia = true) { return this._n.adicionar(lista, checarCheia); }
  /** aplicarFerramenta (fisica-cpu.js:129-152). */
  aplicarFerramenta(f) { if (f && f.tipo && f.tipo !== 'nenhuma') this._n.aplicarFerramenta(f); }
  /** The faucet tool (emitirTorneira, principal.js:301-321) with a Semente in place of
   *  Math.random: ponteiro {x, y, vx, vy}; the semente advances.  Returns {emitidas, aceitas}. */
  emitirTorneira(ponteiro, raio, dt, absorcao = [0, 0, 0], semente = new Semente(1)) {
    const r = this._n.emitirTorneira({ x: ponteiro.x, y: ponteiro.y, vx: ponteiro.vx, vy: ponteiro.vy, raio, absorcao,
      dt, semente: semente.estado });
    semente.estado = r.semente >>> 0;
    return { emitidas: r.emitidas, aceitas: r.aceitas };
  }
  /** avancar (fisica-cpu.js:155-159). */
  avancar(dt, subpassos = 1) { this._n.avancar(dt, subpassos); }
  /** Records the frames run from now on (emitirTorneira and avancar) for retropropagar: the
   *  state is kept every checkpointACada frames and re-run in between (0: every substep kept). */
  gravar(checkpointACada = 10) { this._n.gravar(checkpointACada); }
  /** The walls built on the device from strokes [{x0, y0, x1, y1, raio, apagar}, ...] (Cena.refazer
   *  + atualizarParedes); this.cena is left as it is. */
  definirTracos(tracos) { this._n.definirTracos(Float64Array.from(tracos.flatMap((t) => [t.x0, t.y0, t.x1, t.y1, t.raio, t.apagar ? 1 : 0]))); }
  /** Per stroke, the sum over lattice points of gsdf * d(sdf)/d(x0, y0, x1, y1, raio): a Float64Array n*5. */
  tracosVjp(tracos, gsdf) {
    return this._n.tracosVjp(Float64Array.from(tracos.flatMap((t) => [t.x0, t.y0, t.x1, t.y1, t.raio, t.apagar ? 1 : 0])), gsdf);
  }
  fimGravacao() { this._n.fimGravacao(); }
  /** The render hand-off: the water layer of Render2D.desenhar (render-2d.js:110-162) computed on
   *  the device, {largura: 3 nx, altura: 3 ny, tinta, espuma} with two Uint8ClampedArray RGBA
   *  images, ready for ImageData (tinta multiplied over the panel, espuma drawn over it). */
  renderizarAgua() { return this._n.renderizarAgua(); }
  /** Water fraction of the cell box [i0,i1) x [j0,j1) (mean of min(1, densidade/4)); with
   *  comGradiente, also d(fraction)/d(state) as a Float32Array of n*8 (x y vx vy cux cuy cvx cvy). */
  fracaoAgua(i0, j0, i1, j1, comGradiente = false) { return this._n.fracaoAgua(i0, j0, i1, j1, comGradiente); }
  /** The backward of the recorded frames: from dL/d(final state) (Float32Array n*8) to
   *  {torneira: Float64Array of dL/d(jet vx, vy) per emitirTorneira call, sdf: Float32Array
   *  dL/d(SDF), estadoInicial: Float32Array dL/d(state when gravar was called)}. */
  retropropagar(gradEstado, tol = 1e-6) { return this._n.retropropagar(gradEstado, tol); }
  /** subpasso (fisica-cpu.js:176-189). */
  subpasso(dt) { this._n.subpasso(dt); }
  /** atualizarAr (fisica-cpu.js:162-167). */
  atualizarAr() { this._n.atualizarAr(); }
  /** reiniciarAr (fisica-cpu.js:170-174). */
  reiniciarAr() { this._n.reiniciarAr(); }
  /** area (fisica-cpu.js:192-194). */
  area() { return this._n.area(); }
  /** MotorCPU.limparAgua's effect (motor.js:95): n = 0, counts cleared, air reset. */
  limparAgua() { this._n.limparAgua(); }

  get n() { return this._n.contagem(); }
  set n(m) { if (m === 0) this._n.limparAgua(); else this._n.truncar(m); }
  get vMax() { return this._n.lerVMax(); }
  /** The first n particle records, in the reference's order (a copy). */
  get p() { return this._n.particulas(); }
  particulas() { return this._n.particulas(); }

  get gravidade() { return this._g.slice(); }
  set gravidade(g) {
    this._g = g.slice();
    this._opcoes.gravidade = this._g;
    this._n.opcoes(this._opcoes);
  }

  lerCampo(nome) { return this._n.lerCampo(CAMPO[nome]); }
  get tipo() { return this.lerCampo('tipo'); }
  get rotulo() { return this.lerCampo('rotulo'); }
  get contagemCelula() { return this.lerCampo('contagemCelula'); }
  get solidoCelula() { return this.lerCampo('solidoCelula'); }
  get q() { return this.lerCampo('q'); }

  esperar() { this._n.esperar(); }
  destruir() { this._n.destruir(); }
}

/** Resolves once `pronto()` returns a value other than null, polling between event-loop turns
 *  (the host thread stays free while the device works). */
function aguardar(pronto) {
  return new Promise((resolve, reject) => {
    const volta = () => {
      let v;
      try { v = pronto(); } catch (e) { reject(e); return; }
      if (v !== null) resolve(v); else setImmediate(volta);
    };
    volta();
  });
}

/** MotorCPU's surface (js/motor.js:79-120) on SimulacaoHIP, with MotorGPU's timing (motor.js:
 *  54-74): quadro queues the frame and returns, contagem/particulas/esperar resolve when the
 *  device has reached them, vMax and custoPorSubpasso are those of the last frame whose status
 *  has arrived.  Drawing is the caller's. */
class MotorHIP {
  constructor(canvas, cena, opcoes = {}) {
    this.tipo = 'hip';
    this.nome = 'MI355X';
    this.canvas = canvas;
    this.cena = cena;
    this.sim = new SimulacaoHIP(cena, opcoes);
    this.pendentes = [];
    this.custo = null;
    this.vMaxMedido = null;
    this.subpassosPedidos = 0;
  }

  get h() { return this.cena.h; }
  get maxParticulas() { return Math.ceil(this.sim.nx * this.sim.ny * PARTICULAS_MAX_POR_CELULA); }
  /** Top speed of the last frame whose status has arrived (m/s), or null (MotorGPU.vMax). */
  get vMax() { this._colher(); return this.vMaxMedido; }
  paredesMudaram() { this.sim.atualizarParedes(); }
  adicionar(lista, checar = true) { this.pendentes.push([lista, checar]); }
  limparAgua() { this.sim.limparAgua(); this.pendentes = []; }
  set gravidade(g) { this.sim.gravidade = g; }

  quadro(dt, subpassos, ferramenta) {
    this.sim.aplicarFerramenta(ferramenta);
    // the library cuts each list to the room left under maxParticulas (motor.js:100-104)
    for (const [l, c] of this.pendentes) this.sim.adicionar(l, c);
    this.pendentes = [];
    if (subpassos > 0) {
      this._colher();
      this.sim._n.avancarAssincrono(dt, subpassos);
      this.subpassosPedidos = subpassos;
    }
  }

  /** Takes the last status if it has arrived: vMax and, once per queued frame, its device time
   *  per substep (an EMA as MotorCPU's, motor.js:108-110). */
  _colher() {
    const e = this.sim._n.estado();
    if (e === null) return null;
    this.vMaxMedido = e.vMax;
    if (e.msGpu !== null && this.subpassosPedidos > 0) {
      const ms = e.msGpu / this.subpassosPedidos;
      this.custo = this.custo === null ? ms : this.custo + (ms - this.custo) * 0.2;
      this.subpassosPedidos = 0;
    }
    return e;
  }

  get custoPorSubpasso() { this._colher(); return this.custo; }
  /** The live particle count once the device has finished the queued frames. */
  contagem() {
    this.sim._n.pedirEstado();
    return aguardar(() => { const e = this._colher(); return e === null ? null : e.n; });
  }
  async particulas() { await this.esperar(); return this.sim.particulas(); }
  /** Resolves when the device has finished what was queued (motor.js:72). */
  async esperar() { await this.contagem(); }
  destruir() { this.sim.destruir(); }
}

module.exports = { SimulacaoHIP, MotorHIP, Cena, Semente, semearRetangulo, passoMaximo, CAMPOS, CAMPO };
