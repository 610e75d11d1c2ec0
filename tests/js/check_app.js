'use strict';
// The app-side helpers of the JS host: the per-frame substep plan (Controlador) and the resize
// resampling (reamostrar) with the seeded LCG.  Prints one JSON line of cases for the Python test.
const path = require('path');
const tq = require(path.join(__dirname, '..', '..', 'tanque-de-agua_amd', 'js', 'tanque_hip.js'));

const out = {};
const cena = { h: 0.01, altura: 0.8, largura: 1.42 };
const motor = (v, custo) => ({ vMax: v, custoPorSubpasso: custo });
const c = new tq.Controlador('gpu');
out.semCusto = c.planejar(1 / 20, cena, motor(2.0, null));
out.comCusto = c.planejar(1 / 20, cena, motor(2.0, 4.0));
out.semVmax = c.planejar(1 / 60, cena, motor(null, null));
out.mao = c.planejar(1 / 60, cena, motor(0.5, null), { apertado: true, vx: 3, vy: 4 }, 'mao');
out.passoMao = tq.passoMaximo(cena.h, cena.altura, 5.7, 1 / 60);
out.passo = tq.passoMaximo(cena.h, cena.altura, 2.0, 1 / 20);
c.registrarCusto(10, 2);
c.registrarCusto(20, 2);
out.custoParede = c.custoParede;
const cpu = new tq.Controlador('cpu');
out.cpu = cpu.planejar(1 / 20, cena, motor(2.0, null));

const lista = new Float32Array(3 * tq.CAMPOS);
const xs = [[0.5, 0.4], [0.015, 0.4], [0.7, 0.3]];
xs.forEach(([x, y], k) => { lista[k * tq.CAMPOS] = x; lista[k * tq.CAMPOS + 1] = y; lista[k * tq.CAMPOS + 2] = k + 1; });
let s = new tq.Semente(7);
out.igual = Array.from(tq.reamostrar(lista, 0.01, 0.01, 1.42, 0.8, () => s.proximo()));
out.igualSemente = s.estado;
s = new tq.Semente(7);
out.quatro = Array.from(tq.reamostrar(lista, 0.02, 0.01, 1.42, 0.8, () => s.proximo()));
out.quatroSemente = s.estado;
const grande = new Float32Array(4000 * tq.CAMPOS);
for (let k = 0; k < 4000; k++) { grande[k * tq.CAMPOS] = 0.2 + (k % 80) * 0.01; grande[k * tq.CAMPOS + 1] = 0.1 + Math.floor(k / 80) * 0.01; }
s = new tq.Semente(11);
out.fracionario = tq.reamostrar(grande, 0.015, 0.01, 1.42, 0.8, () => s.proximo()).length / tq.CAMPOS;
process.stdout.write(JSON.stringify(out) + '\n');
