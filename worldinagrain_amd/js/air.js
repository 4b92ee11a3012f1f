// The air's projection for the JavaScript host: what the reference's host
// keeps around the solve (js/sim.js:36,44,738 airCycles; :1916-1920 its
// update from the iterations a step ran), over wig_air_project through the
// wig.node addon in place of encodeAir's solve passes (js/sim.js:2264-2289).
// The buffers are the reference's: airProjection ("mg", every level, two
// vec4s a cell), airCg ("cg"), airFastest's words ("status"), and this
// library's "mean" (each cell's zone's meanExpansion).  CommonJS, Node 12.
"use strict";

const path = require("path");
const addon = require(path.join(__dirname, "wig.node"));

const AIR_CYCLES = 64;        // js/sim.js:36
const AIR_CYCLES_SPARE = 8;   // js/sim.js:44
const PART = { mg: 0, cg: 1, status: 2, mean: 3 };

class AirProjection {
    constructor(width, height) {
        this.handle = addon.airCreate(width, height);
        this.layout = addon.airLayout(this.handle);
        this.width = width;
        this.height = height;
        this.airCycles = AIR_CYCLES;
    }

    upload(part, bytes) { addon.airUpload(this.handle, PART[part], bytes); }

    download(part) {
        const out = new Uint8Array(this.layout.partBytes[PART[part]]);
        addon.airDownload(this.handle, PART[part], out);
        return out;
    }

    // One step's solve with the iterations the host gives it, then the next
    // step's count from what this one ran (the report's AIR_MOST_CYCLES).
    project() {
        const given = this.airCycles;
        addon.airProject(this.handle, given);
        const status = new Uint32Array(this.download("status").buffer);
        const most = status[4];
        this.airCycles = most >= given ? AIR_CYCLES : Math.min(2 * most + AIR_CYCLES_SPARE, AIR_CYCLES);
        return most;
    }

    destroy() {
        addon.airDestroy(this.handle);
        this.handle = null;
    }
}

module.exports = { AirProjection, AIR_CYCLES, PART };
