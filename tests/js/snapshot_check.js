// The snapshot container from the JavaScript side, for the Python tests:
//   numbers <f64 file>          JSON.stringify of the numbers (the engine's own printing)
//   read <snapshot file>        the header and each part's sha256
//   pack <header json> <dir>    the uncompressed container of the header's parts (<dir>/<name>.bin) to stdout
"use strict";
const fs = require("fs");
const path = require("path");
const crypto = require("crypto");
const snap = require(path.join(__dirname, "..", "..", "worldinagrain_amd", "js", "snapshot.js"));

const [mode, a, b] = process.argv.slice(2);
if (mode === "numbers") {
    const buf = fs.readFileSync(a);
    const v = new Float64Array(buf.buffer, buf.byteOffset, buf.byteLength / 8);
    process.stdout.write(JSON.stringify({ v: Array.from(v) }));
} else if (mode === "read") {
    const { header, parts } = snap.readSnapshot(fs.readFileSync(a));
    const digests = {};
    for (const [name, bytes] of parts) digests[name] = crypto.createHash("sha256").update(bytes).digest("hex");
    process.stdout.write(JSON.stringify({ header, digests }));
} else if (mode === "pack") {
    const header = JSON.parse(fs.readFileSync(a, "utf8"));
    const parts = new Map(header.parts.map((p) => [p.name, fs.readFileSync(path.join(b, p.name + ".bin"))]));
    process.stdout.write(snap.packSnapshot(header, parts));
} else {
    throw new Error("mode: numbers, read or pack");
}
