'use strict';
// Drives the adjoint through the Node-API addon as a JavaScript optimiser would: the funnel
// scene (strokes given as JSON), the faucet every frame and avancar with the adaptive substep
// bound while recording, then the water fraction of a box, the loss (fraction - 0.5)^2 and
// retropropagar.  Prints one JSON line: fraction, loss, the summed jet gradient and the SDF
// gradient (base64 of a Float32Array).
// Usage: node tests/js/run_funnel_adjoint.js <json: {nx, ny, h, strokes, faucet, frames, dt, box, every}>
const path = require('path');
const tq = require(path.join(__dirname, '..', '..', 'tanque-de-agua_amd', 'js', 'tanque_hip.js'));

const cfg = JSON.parse(process.argv[2]);
const cena = new tq.Cena(cfg.nx, cfg.ny, cfg.h);
for (const s of cfg.strokes) cena.tracar(s[0], s[1], s[2], s[3], s[4], s[5]);
const sim = new tq.SimulacaoHIP(cena, { iteracoes: 200, tol: 1e-6 });
const f = cfg.faucet;
const semente = new tq.Semente(f.seed);
sim.gravar(cfg.every);
for (let k = 0; k < cfg.frames; k++) {
  sim.emitirTorneira({ x: f.x, y: f.y, vx: f.pvx, vy: f.pvy }, f.raio, cfg.dt, [0, 0, 0], semente);
  const sub = Math.ceil(cfg.dt / tq.passoMaximo(cfg.h, cena.altura, sim.vMax, cfg.dt));
  sim.avancar(cfg.dt, sub);
}
const b = cfg.box;
const r = sim.fracaoAgua(b[0], b[1], b[2], b[3], true);
const w = 2 * (r.fracao - 0.5);
const g = new Float32Array(r.gradiente.length);
for (let i = 0; i < g.length; i++) g[i] = w * r.gradiente[i];
const back = sim.retropropagar(g, 1e-6);
let gx = 0, gy = 0;
for (let c = 0; c < back.torneira.length; c += 2) { gx += back.torneira[c]; gy += back.torneira[c + 1]; }
const out = { fracao: r.fracao, perda: (r.fracao - 0.5) ** 2, jato: [gx, gy], chamadas: back.torneira.length / 2,
              sdf: Buffer.from(back.sdf.buffer, back.sdf.byteOffset, back.sdf.byteLength).toString('base64') };
sim.fimGravacao();
sim.destruir();
process.stdout.write(JSON.stringify(out) + '\n');
