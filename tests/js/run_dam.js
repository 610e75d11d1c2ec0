'use strict';
// Drives the library through the Node-API addon the way the reference's tests drive
// SimulacaoCPU (testes/fisica.test.js:12-29): a dam-break tank stepped with the adaptive
// substep bound; prints one JSON line with the particle records (base64 of the Float32Array).
// Usage: node tests/js/run_dam.js <nx> <ny> <h> <seconds> [host-only | render]
// (render: also the water layer of the end state, base64 of the two RGBA images)
const path = require('path');
const tq = require(path.join(__dirname, '..', '..', 'tanque-de-agua_amd', 'js', 'tanque_hip.js'));

const [nx, ny, h, segundos] = process.argv.slice(2, 6).map(Number);
const hostOnly = process.argv[6] === 'host-only';
const render = process.argv[6] === 'render';
const MOLDURA = 2;
const m = MOLDURA * h;
const semente = new tq.Semente(9);
const agua = tq.semearRetangulo(m, m, 0.3, 0.42, h, [0, 0, 0], semente);
const out = { n_seeded: agua.length / tq.CAMPOS, passo: tq.passoMaximo(h, ny * h, null, 1 / 60),
              semente: semente.estado };
async function main() {
  if (hostOnly) {
    out.particulas = Buffer.from(agua.buffer).toString('base64');
    return;
  }
  const cena = new tq.Cena(nx, ny, h);
  const motor = new tq.MotorHIP(null, cena, {});
  motor.adicionar(agua, false);
  const quadro = 1 / 60;
  const quadros = Math.floor(segundos / quadro + 0.5);
  for (let f = 0; f < quadros; f++) {
    // the plan reads the previous frame's top speed: wait for its status (the app plans with
    // whatever arrived, as with MotorGPU)
    await motor.esperar();
    const sub = Math.ceil(quadro / tq.passoMaximo(h, cena.altura, motor.vMax, quadro));
    motor.quadro(quadro, sub, null);
  }
  await motor.esperar();
  const P = motor.sim.particulas();
  out.n = motor.sim.n;
  out.vMax = motor.vMax;
  out.area = motor.sim.area();
  out.custoPorSubpasso = motor.custoPorSubpasso;
  out.particulas = Buffer.from(P.buffer, P.byteOffset, P.byteLength).toString('base64');
  if (render) {
    const img = motor.sim.renderizarAgua();
    out.largura = img.largura;
    out.altura = img.altura;
    out.tinta = Buffer.from(img.tinta.buffer).toString('base64');
    out.espuma = Buffer.from(img.espuma.buffer).toString('base64');
  }
  motor.destruir();
}
main().then(() => process.stdout.write(JSON.stringify(out) + '\n'), (e) => { console.error(e); process.exit(1); });
