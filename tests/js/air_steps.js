// Runs the air's projection from the JavaScript host: reads the grid's
// buffers (mg, cg, status, mean) from dir, steps AirProjection `steps`
// times (the same level 0 each step, the potentials warm), writes the
// buffers back and the iterations of each step.
"use strict";
const fs = require("fs");
const path = require("path");
const { AirProjection } = require(path.join(__dirname, "..", "..", "worldinagrain_amd", "js", "air.js"));

const [W, H, steps, dir] = [Number(process.argv[2]), Number(process.argv[3]), Number(process.argv[4]), process.argv[5]];
const air = new AirProjection(W, H);
for (const part of ["mg", "cg", "status", "mean"]) air.upload(part, fs.readFileSync(path.join(dir, part + ".bin")));
const level0 = fs.readFileSync(path.join(dir, "mg.bin")).subarray(0, 32 * W * H);
const its = [];
for (let s = 0; s < steps; s++) {
    if (s > 0) {
        const mg = Buffer.from(air.download("mg"));
        level0.copy(mg, 0);
        air.upload("mg", mg);
    }
    its.push(air.project());
}
for (const part of ["mg", "cg", "status"]) fs.writeFileSync(path.join(dir, part + ".out"), Buffer.from(air.download(part)));
fs.writeFileSync(path.join(dir, "its.json"), JSON.stringify(its));
air.destroy();
