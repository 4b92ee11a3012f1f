// Steps a dam break through the JavaScript host (liquid.js over wig.node)
// and prints, as JSON, each step's FLUID_STATS and a digest of the cells,
// counters and field after the last step.
"use strict";
const path = require("path");
const crypto = require("crypto");
const { LiquidSimulation, PART } = require(path.join(__dirname, "..", "..", "worldinagrain_amd", "js", "liquid.js"));

const [W, H, steps] = process.argv.slice(2).map(Number);
const cellsPath = process.argv[5];
const fs = require("fs");
const cells = fs.readFileSync(cellsPath);
const sim = new LiquidSimulation(W, H);
sim.upload(PART.cells, new Uint8Array(cells.buffer, cells.byteOffset, cells.byteLength));
const stats = [];
for (let s = 0; s < steps; s++) {
    sim.encodeFluid();
    stats.push(Array.from(sim.readReport()));
}
const digest = {};
for (const name of ["cells", "counters", "field"]) {
    digest[name] = crypto.createHash("sha256").update(Buffer.from(sim.download(PART[name]))).digest("hex");
}
sim.destroy();
console.log(JSON.stringify({ stats, digest }));
