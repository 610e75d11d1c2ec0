// bench_oracle.js -- times the oracle (the C++ restatement of the reference CPU engine) from
// Node through oracle/oracle_node.node, on a bounded sample of a workload written by bench.py:
//   node oracle/bench_oracle.js workload.json
// workload.json: { nx, ny, h, strokes: [[x0, y0, x1, y1, raio, apagar], ...], particles_file (raw
// little-endian float32, 12 per particle), dt_frame, frames, budget_s, tol, maxit }.
// Frames are split as the reference's tests split them (fisica.test.js:12-19):
// ceil(dt / passoMaximo(h, height, vMax, dt)) substeps.  Prints one JSON line.
'use strict';
const fs = require('fs');
const path = require('path');

const addon = require(path.join(__dirname, 'oracle_node.node'));
const w = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));

const cena = addon.cena(w.nx, w.ny, w.h);
for (const s of w.strokes) addon.tracar(cena, s[0], s[1], s[2], s[3], s[4], !!s[5]);
const sim = addon.simulacao(cena, w.tol, w.maxit);
const buf = fs.readFileSync(w.particles_file);
const lista = new Float32Array(buf.buffer.slice(buf.byteOffset, buf.byteOffset + buf.byteLength));
const n = addon.adicionar(sim, lista, false);
const altura = w.ny * w.h;
const t0 = process.hrtime.bigint();
let frames = 0;
let substeps = 0;
let seconds = 0;
while (frames < w.frames && seconds < w.budget_s) {
  const sub = Math.ceil(w.dt_frame / addon.passoMaximo(w.h, altura, addon.vmax(sim), w.dt_frame));
  addon.avancar(sim, w.dt_frame, sub);
  substeps += sub;
  frames += 1;
  seconds = Number(process.hrtime.bigint() - t0) / 1e9;
}
process.stdout.write(JSON.stringify({ particles: n, frames, substeps, seconds }) + '\n');
