'use strict';
// MotorHIP's asynchronous frame (MotorGPU's timing, motor.js:54-74): frames queued without
// waiting, then the count, vMax and cost read back when the device gets there, against the
// blocking reads of the same state; and the particle cap of quadro (motor.js:100-104).  Prints
// one JSON line for the Python test.
const path = require('path');
const tq = require(path.join(__dirname, '..', '..', 'tanque-de-agua_amd', 'js', 'tanque_hip.js'));

async function main() {
  const out = {};
  const nx = 240, ny = 135, h = 0.8 / 135;
  const m = 2 * h;
  const cena = new tq.Cena(nx, ny, h);
  const motor = new tq.MotorHIP(null, cena, {});
  motor.adicionar(tq.semearRetangulo(m, m, 0.45, 0.5, h, [0, 0, 0], new tq.Semente(4)), false);
  out.vMaxAntes = motor.vMax;  // no frame yet
  const quadros = 30, sub = 3;
  const t0 = process.hrtime.bigint();
  for (let f = 0; f < quadros; f++) motor.quadro(1 / 120, sub, f === 10 ? { tipo: 'mao', x: 0.2, y: 0.2, raio: 0.1, vx: 2, vy: 0 } : null);
  const t1 = process.hrtime.bigint();
  out.prontoLogo = motor.sim._n.estado() !== null;
  out.n = await motor.contagem();
  const t2 = process.hrtime.bigint();
  out.msFila = Number(t1 - t0) / 1e6;
  out.msAteContagem = Number(t2 - t0) / 1e6;
  out.vMax = motor.vMax;
  out.nBloqueante = motor.sim.n;           // tq_count: waits for the stream
  out.vMaxBloqueante = motor.sim.vMax;     // tq_vmax: waits for the stream
  out.custo = motor.custoPorSubpasso;
  // the cap: lists beyond maxParticulas are cut to the room left (motor.js:100-104)
  out.teto = motor.maxParticulas;
  const d = 4 * h;  // 4 cells inside the frame: adicionar's wall rule passes every particle
  for (let k = 0; k < 3; k++) motor.adicionar(tq.semearRetangulo(d, d, 1.42 - d, 0.8 - d, h, [0, 0, 0], new tq.Semente(10 + k)), false);
  motor.quadro(1 / 120, 1, null);
  out.nNoTeto = await motor.contagem();
  motor.adicionar(tq.semearRetangulo(m, m, 0.3, 0.3, h, [0, 0, 0], new tq.Semente(20)), false);
  motor.quadro(1 / 120, 1, null);
  out.nDepois = await motor.contagem();
  const P = await motor.particulas();
  out.nParticulas = P.length / tq.CAMPOS;
  motor.destruir();
  return out;
}
main().then((o) => process.stdout.write(JSON.stringify(o) + '\n'), (e) => { console.error(e); process.exit(1); });
