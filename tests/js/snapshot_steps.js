// Steps a dam break through the JavaScript host, saving a snapshot after the
// first `steps` steps and another after as many more (for the Python tests).
"use strict";
const fs = require("fs");
const path = require("path");
const { LiquidSimulation, PART } = require(path.join(__dirname, "..", "..", "worldinagrain_amd", "js", "liquid.js"));
const snap = require(path.join(__dirname, "..", "..", "worldinagrain_amd", "js", "snapshot.js"));

const [W, H, steps] = process.argv.slice(2, 5).map(Number);
const [cellsPath, outA, outB] = process.argv.slice(5);
const cells = fs.readFileSync(cellsPath);
const sim = new LiquidSimulation(W, H);
sim.upload(PART.cells, new Uint8Array(cells.buffer, cells.byteOffset, cells.byteLength));
sim.step(steps);
fs.writeFileSync(outA, snap.saveSnapshot(sim, 1));
// (the snapshot read back into a second world of the JavaScript host, which
// steps on as the first)
const again = new LiquidSimulation(W, H);
snap.restoreSnapshot(again, snap.readSnapshot(fs.readFileSync(outA)));
sim.step(steps);
again.step(steps);
fs.writeFileSync(outB, snap.saveSnapshot(sim, 1));
const a = Buffer.from(sim.download(PART.field)), b = Buffer.from(again.download(PART.field));
sim.destroy();
again.destroy();
process.stdout.write(JSON.stringify({ fieldSame: a.equals(b) }));
