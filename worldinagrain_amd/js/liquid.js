// The liquid's part of worldinagrain's host Simulation (js/sim.js), over the
// HIP liquid step: what the reference keeps and decides around
// encodeFluid (js/sim.js:1386-1397 seed and random, :1893-1901 the report's
// fluidStats and fluidCycles, :1972-2001 writeSlot, :2466-2612 encodeFluid),
// calling libwig.so through the wig.node addon instead of encoding WGSL
// passes.  CommonJS, for Node 12.
"use strict";

const path = require("path");
const addon = require(path.join(__dirname, "wig.node"));

const FLUID_CYCLES = 16;        // js/sim.js:63
const FLUID_CYCLES_SPARE = 2;   // js/sim.js:70
const LIQUID_TOLERANCE = 1e-2;  // fluid.wgsl:2221
const GRAVITY = 9.81 / 3600 / 0.01;   // js/elements.js:30, water's weight
const CAPILLARY_LENGTH = 0.0027;       // js/shared.js:285-286
const SURFACE_MIN_LENGTH = 0.3;

const PART = { cells: 0, particles0: 1, particles1: 2, counters: 3, field: 4, blocks: 5, fluidCg: 6, solve: 7, air0: 8,
    airSource: 9, occupants: 10, gravity: 11, bodyMotion: 12, bodySwept: 13, bodyPushes: 14 };
const STAT = { ACTIVE: 0, RETIRED: 1, VISCOUS: 2, PARTICLES: 3, RESIDUAL: 4, CYCLES: 7, MOST_CYCLES: 8 };
const TOL_ABS_MAX = 0;
const TOL_REL_L2 = 1;
const CHUNK_AUTO = 0xffffffff;
const KIND = { SOLID: 1, POWDER: 2, LIQUID: 3, GAS: 4 };

// The wind's drag coefficient along a liquid surface, (KARMAN / ln(max(dx /
// 2 / z0, 2)))^2 in f32 (fluid.wgsl:1884-1895; KARMAN js/shared.js:289), as
// the Python host works it out, so both hand the library the same float.
function surfaceDragCoefficient(dx) {
    const f = Math.fround;
    const r = f(f(0.41) / f(Math.log(Math.max(f(f(f(0.5) * f(dx)) / f(1e-4)), 2))));
    return f(r * r);
}

// elems[t].kind for the ids of js/elements.js's table: AIR 0, WALL 1,
// SAND 2, WATER 3, STONE 4, GRAVEL 5.
function elementKinds() {
    const k = new Uint8Array(256);
    k.set([KIND.GAS, KIND.SOLID, KIND.POWDER, KIND.LIQUID, KIND.SOLID, KIND.POWDER]);
    return k;
}

class LiquidSimulation {
    // options: seed, dx (metres a cell), earthG (cells/step^2), dt (seconds
    // a step), frame, and the solve's settings (solver: tolMode, tol,
    // maxCycles, adaptive, preRelax, postRelax, sweepPairs, coarseSweeps;
    // the reference's own by default).
    constructor(width, height, options = {}) {
        this.width = width;
        this.height = height;
        this.world = addon.create(width, height, options.coarseLevels || 0);
        addon.setKinds(this.world, options.kinds || elementKinds());
        // the elements' columns (viscosity, its temperature law, tension,
        // wetting, breaks): 256 rows of 6 floats, zeros unless given
        if (options.elementProps) addon.setElementProps(this.world, options.elementProps);
        this.layout = addon.layout(this.world);
        this.rngState = (options.seed === undefined ? 0x5eed : options.seed) >>> 0;
        this.metresPerCell = options.dx === undefined ? 0.01 : options.dx;
        this.earthG = options.earthG === undefined ? GRAVITY : options.earthG;
        this.dt = options.dt === undefined ? 1 / 60 : options.dt;
        this.frameNo = options.frame === undefined ? 1 : options.frame;
        this.air = !!options.air;      // the air coupled in (air0 read, airSource added to)
        this.occupants = !!options.occupants;   // people's bodies (the occupants part) as moving obstacles
        this.gravity = !!options.gravity;       // the gravity pass's field (the gravity part): pull and down
        this.bodies = !!options.bodies;         // free and held bodies (bodyMotion, bodySwept; bodyPushes scratch)
        this.bodiesRun = false;                 // bodiesRun(): the caller's body pass sets it each step
        const s = options.solver || {};
        this.solver = {
            tolMode: s.tolMode === undefined ? TOL_ABS_MAX : s.tolMode,
            tol: s.tol === undefined ? LIQUID_TOLERANCE : s.tol,
            maxCycles: s.maxCycles || FLUID_CYCLES,
            adaptive: s.adaptive === undefined ? true : s.adaptive,
            preRelax: s.preRelax || 1,
            postRelax: s.postRelax || 1,
            sweepPairs: s.sweepPairs || 4,
            coarseSweeps: s.coarseSweeps || 40,
            // this build's chunked encoding: 0 encodes maxCycles iterations as
            // the reference does; CHUNK_AUTO as many as the last solve ran,
            // then `chunk` more at a time until it converges
            firstChunk: s.firstChunk || 0,
            chunk: s.chunk || 2,
        };
        this.shiftedFirst = false;
        this.fluidStats = { active: 1, viscous: 0 };
        this.fluidCycles = this.solver.maxCycles;
        this.lastStats = null;
    }

    destroy() {
        addon.destroy(this.world);
    }

    // Simulation.seed / random: mulberry32 (Tommy Ettinger's, public domain)
    seed(n) {
        this.rngState = n >>> 0;
    }

    random() {
        this.rngState = (this.rngState + 0x6d2b79f5) >>> 0;
        let t = this.rngState;
        t = Math.imul(t ^ (t >>> 15), t | 1);
        t ^= t + Math.imul(t ^ (t >>> 7), t | 61);
        return (t ^ (t >>> 14)) >>> 0;
    }

    // The Params slot as writeSlot fills the words the liquid reads.
    writeSlot() {
        const buf = new ArrayBuffer(80);
        const u = new Uint32Array(buf);
        const f = new Float32Array(buf);
        u[0] = this.width;
        u[1] = this.height;
        u[2] = this.frameNo;
        u[3] = this.random();
        f[7] = this.metresPerCell;
        f[12] = this.earthG;
        f[13] = this.dt;
        return buf;
    }

    // Whether a cell is small enough for surface tension to show
    // (Simulation.surfaceTension, js/sim.js:2449-2454).
    surfaceTension() {
        return CAPILLARY_LENGTH / this.metresPerCell >= SURFACE_MIN_LENGTH;
    }

    // What encodeFluid decides: idle or full, the iterations to encode, the
    // relaxation's starting tiling (flipped each step), the solve's settings.
    stepControl() {
        this.shiftedFirst = !this.shiftedFirst;
        const buf = new ArrayBuffer(88);
        const u = new Uint32Array(buf);
        const f = new Float32Array(buf);
        // bodiesRun() makes the step full and hands the bodies what the
        // liquid did to them (js/sim.js:2479, 2596-2599)
        const run = this.bodies && this.bodiesRun;
        u[0] = this.fluidStats.active || run ? 1 : 0;
        u[3] = run ? 1 : 0;
        // the surface pass, and viscosity's sweeps with it (js/sim.js:2496-2509)
        u[2] = this.surfaceTension() ? 1 : 0;
        u[1] = this.fluidStats.viscous || u[2] ? 1 : 0;
        u[4] = this.solver.adaptive ? this.fluidCycles : this.solver.maxCycles;
        u[5] = this.shiftedFirst ? 1 : 0;
        u[6] = this.solver.tolMode;
        f[7] = this.solver.tol;
        u[8] = this.solver.preRelax;
        u[9] = this.solver.postRelax;
        u[10] = this.solver.sweepPairs;
        u[11] = this.solver.coarseSweeps;
        u[12] = this.solver.firstChunk >>> 0;
        u[13] = this.solver.chunk;
        u[16] = this.solver.residualEvery || 0;
        u[17] = this.air ? 1 : 0;
        f[18] = this.air ? surfaceDragCoefficient(this.metresPerCell) : 0;
        u[19] = this.occupants ? 1 : 0;
        u[20] = this.gravity ? 1 : 0;
        u[21] = this.bodies ? 1 : 0;
        return buf;
    }

    // One step's liquid (encodeFluid), then the report taken in.
    encodeFluid() {
        const status = addon.fluidStep(this.world, this.writeSlot(), this.stepControl());
        this.frameNo++;
        return status;
    }

    // The report asked for now and taken in when it lands (readReport's
    // mapAsync, js/sim.js:1882-1901): the last step's, so the host never
    // waits for the step it has just enqueued.
    readReportLate() {
        const last = this.pendingReport;
        this.pendingReport = addon.reportAsync(this.world);
        return last === undefined ? null : this.takeReport(addon.reportTake(this.world, last, true));
    }

    readReport() {
        return this.takeReport(addon.report(this.world));
    }

    takeReport(u) {
        this.lastStats = u;
        this.fluidStats = { active: u[STAT.ACTIVE] + u[STAT.PARTICLES], viscous: u[STAT.VISCOUS] };
        const most = u[STAT.MOST_CYCLES];
        const cap = this.solver.maxCycles;
        this.fluidCycles = most >= this.fluidCycles ? cap : Math.min(most + FLUID_CYCLES_SPARE, cap);
        return u;
    }

    step(n = 1) {
        for (let i = 0; i < n; i++) {
            this.encodeFluid();
            this.readReport();
        }
    }

    upload(part, bytes) {
        addon.upload(this.world, part, bytes);
    }

    download(part) {
        const out = new ArrayBuffer(this.layout.partBytes[part]);
        addon.download(this.world, part, out);
        return out;
    }

    solveStats() {
        return addon.solveStats(this.world);
    }
}

module.exports = { LiquidSimulation, PART, STAT, KIND, TOL_ABS_MAX, TOL_REL_L2, CHUNK_AUTO, GRAVITY, elementKinds, deviceOk: addon.deviceOk };
