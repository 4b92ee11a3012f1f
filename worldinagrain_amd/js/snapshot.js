// World snapshots for the JavaScript host, in the reference's container
// (js/snapshot.js:1-19,97-192,512-545): gzip over the magic "SANDSNAP", the
// header's length (u32, little-endian), the header (JSON.stringify's), and
// each part's bytes in the header's order, each padded to 4 with zeros.  The
// parts are the liquid's buffers by the reference's names (cells as
// "cells<cur>", particles0, particles1, counters, field, blocks, fluidCg)
// and this library's "solve"; the header's state holds snapshotState's
// liquid fields (js/sim.js:1445-1465).  The same files as the Python host's
// worldinagrain_amd/snapshot.py.  CommonJS, for Node 12 (zlib's gzip in
// place of the browser's CompressionStream).
"use strict";

const zlib = require("zlib");

const MAGIC = "SANDSNAP";
const FORMAT = 1;
// the element ids of the liquid step's table (js/elements.js:416-457)
const ELEMENT_IDS = ["AIR", "WALL", "SAND", "WATER", "STONE", "GRAVEL"];
// parts this library keeps that the reference's world does not: loaded
// empty when missing (as js/snapshot.js:31-43's ADDED parts do)
const ADDED = new Set(["solve"]);
const NAMES = ["cells", "particles0", "particles1", "counters", "field", "blocks", "fluidCg", "solve"];

const align = (n) => Math.ceil(n / 4) * 4;
const partName = (p, cur = 0) => (p === 0 ? `cells${cur}` : NAMES[p]);   // (part 0: the cells)

// The uncompressed bytes (saveSnapshot's packing, js/snapshot.js:139-157).
function packSnapshot(header, parts) {
    const head = Buffer.from(JSON.stringify(header), "utf8");
    const start = MAGIC.length + 4 + align(head.length);
    const out = Buffer.alloc(start + header.parts.reduce((n, p) => n + align(p.bytes), 0));
    out.write(MAGIC, 0, "latin1");
    out.writeUInt32LE(head.length, MAGIC.length);
    head.copy(out, MAGIC.length + 4);
    let at = start;
    for (const p of header.parts) {
        const bytes = Buffer.from(parts.get(p.name));
        if (bytes.length !== p.bytes) {
            throw new Error(`part ${p.name}: ${bytes.length} bytes, the header says ${p.bytes}`);
        }
        bytes.copy(out, at);
        at += align(p.bytes);
    }
    return out;
}

// The world as a snapshot file's bytes (saveSnapshot).
function saveSnapshot(sim, level = 6) {
    const parts = new Map();
    const items = [];
    for (let p = 0; p < NAMES.length; p++) {
        const bytes = new Uint8Array(sim.download(p));
        parts.set(partName(p), bytes);
        items.push({ name: partName(p), bytes: bytes.length });
    }
    const header = {
        format: FORMAT, width: sim.width, height: sim.height, elements: ELEMENT_IDS.slice(),
        state: {
            cur: 0, frameNo: sim.frameNo, shiftedFirst: sim.shiftedFirst, rng: sim.rngState,
            metresPerCell: sim.metresPerCell, earthGravity: sim.earthG,
            fluidStats: { active: sim.fluidStats.active, viscous: sim.fluidStats.viscous },
            fluidCycles: sim.fluidCycles, dt: sim.dt,
        },
        parts: items,
    };
    return zlib.gzipSync(packSnapshot(header, parts), { level });
}

// A snapshot unpacked: { header, parts }, parts a Map from name to bytes
// (readSnapshot, js/snapshot.js:166-192); the same refusals.
function readSnapshot(bytes) {
    let raw;
    try {
        raw = zlib.gunzipSync(bytes);
    } catch (e) {
        throw new Error("not a snapshot");
    }
    if (raw.length < MAGIC.length + 4 || raw.toString("latin1", 0, MAGIC.length) !== MAGIC) {
        throw new Error("not a snapshot");
    }
    const length = raw.readUInt32LE(MAGIC.length);
    const header = JSON.parse(raw.toString("utf8", MAGIC.length + 4, MAGIC.length + 4 + length));
    if (header.format !== FORMAT) {
        throw new Error(`saved in snapshot format ${header.format}; this version reads format ${FORMAT}`);
    }
    const parts = new Map();
    let at = MAGIC.length + 4 + align(length);
    for (const part of header.parts) {
        if (at + part.bytes > raw.length) {
            throw new Error("the snapshot is cut short");
        }
        parts.set(part.name, raw.subarray(at, at + part.bytes));
        at += align(part.bytes);
    }
    return { header, parts };
}

// Loads a snapshot (readSnapshot's) into sim, a world of its size
// (restoreSnapshot, js/snapshot.js:512-545): the same elements, each part
// this world's size (one of ADDED may be missing: it loads empty), then the
// state.  Throws, with nothing changed, if any part doesn't fit.
function restoreSnapshot(sim, { header, parts }) {
    if (header.width !== sim.width || header.height !== sim.height) {
        throw new Error(`the snapshot is of a ${header.width}×${header.height} world, this one is ${sim.width}×${sim.height}`);
    }
    if (header.elements.join() !== ELEMENT_IDS.join()) {
        throw new Error("saved with a different set of elements");
    }
    const cur = header.state.cur || 0;
    const load = [];
    const misfits = [];
    for (let p = 0; p < NAMES.length; p++) {
        const name = partName(p, cur);
        const size = sim.layout.partBytes[p];
        const saved = parts.get(name);
        if (saved === undefined && ADDED.has(partName(p))) {
            load.push([p, new Uint8Array(size)]);
        } else if (saved === undefined || saved.length !== size) {
            misfits.push(name);
        } else {
            load.push([p, new Uint8Array(saved.buffer, saved.byteOffset, saved.length)]);
        }
    }
    if (misfits.length) {
        throw new Error(`saved by a version that lays out ${misfits.join(", ")} differently`);
    }
    for (const [p, bytes] of load) {
        sim.upload(p, bytes);
    }
    const s = header.state;
    sim.frameNo = s.frameNo;
    sim.shiftedFirst = s.shiftedFirst;
    sim.rngState = s.rng >>> 0;
    sim.metresPerCell = s.metresPerCell;
    sim.earthG = s.earthGravity;
    sim.fluidStats = { active: s.fluidStats.active, viscous: s.fluidStats.viscous };
    sim.fluidCycles = s.fluidCycles;
    if (s.dt !== undefined) {
        sim.dt = s.dt;
    }
}

module.exports = { MAGIC, FORMAT, ELEMENT_IDS, packSnapshot, saveSnapshot, readSnapshot, restoreSnapshot };
