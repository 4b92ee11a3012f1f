// Steps a dam break through the JavaScript host (liquid.js over wig.node)
// and prints, as JSON, each step's FLUID_STATS and a digest of the cells,
// counters and field after the last step.  With a sixth argument "late",
// each report is asked for without waiting and taken in a step later
// (readReportLate), the last one after the loop.
"use strict";
const path = require("path");
const crypto = require("crypto");
const { LiquidSimulation, PART } = require(path.join(__dirname, "..", "..", "worldinagrain_amd", "js", "liquid.js"));
const addon = require(path.join(__dirname, "..", "..", "worldinagrain_amd", "js", "wig.node"));
const addonTake = (sim) => addon.reportTake(sim.world, sim.pendingReport, true);

const [W, H, steps] = process.argv.slice(2).map(Number);
const cellsPath = process.argv[5];
const fs = require("fs");
const cells = fs.readFileSync(cellsPath);
const sim = new LiquidSimulation(W, H);
sim.upload(PART.cells, new Uint8Array(cells.buffer, cells.byteOffset, cells.byteLength));
const late = process.argv[6] === "late";
const stats = [];
for (let s = 0; s < steps; s++) {
    sim.encodeFluid();
    if (late) {
        const u = sim.readReportLate();
        if (u) stats.push(Array.from(u));
    } else {
        stats.push(Array.from(sim.readReport()));
    }
}
if (late) stats.push(Array.from(sim.takeReport(addonTake(sim))));
const digest = {};
for (const name of ["cells", "counters", "field"]) {
    digest[name] = crypto.createHash("sha256").update(Buffer.from(sim.download(PART[name]))).digest("hex");
}
sim.destroy();
console.log(JSON.stringify({ stats, digest }));
