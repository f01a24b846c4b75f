'use strict';
// createProbeField: the probe-field controller of js/gi_probes.js:4548-4844, driving the
// MI355X library through the N-API addon. Same method names, the same argument
// normalisation for every knob that reaches the probe update, and the same async
// tick() with an in-flight guard. Knobs that only affect the renderer's sampling
// (reflection, intensity, Chebyshev strength, normal detail) are kept as state so
// callers keep working; the renderer is outside this build.
const { loadAddon, STATUS, BUF, BLEND, TRAVERSAL, REFLECTION } = require('./addon');
const { buildScene, worldMatrix, sceneBounds } = require('./scene_ingest');

// js/gi_probes.js:53-191 constants that shape the update
const OCT_RES = 6;
const TILE = 8;
const RAYS_PER_PROBE_DEFAULT = 64;
const RAYS_MIN = 32, RAYS_MAX = 256;
const TARGET_PROBES_LONG_AXIS = 12;
const MAX_PROBES_PER_AXIS = 32;
const RAYS_PER_TICK = 98304;
const RAYS_PER_TICK_MIN = 2048;
const FIELD_CLUSTERED = 2;  // SBGI_FIELD_CLUSTERED
const MAX_PROBES_PER_TICK = 2048;
const JITTER_HYSTERESIS_DEFAULTS = Object.freeze({ gated: 0.60, montecarlo: 0.90 });
const GOLDEN_RATIO_FRACT = 0.6180339887498949;

function clamp(v, lo, hi) { return Math.min(hi, Math.max(lo, v)); }
function fract(x) { return x - Math.floor(x); }

// computeGridResolution (js/gi_probes.js:938-943)
function computeGridResolution(size, targetLongAxis) {
    const longest = Math.max(size[0], size[1], size[2], 1e-3);
    const spacing = longest / Math.max(1, targetLongAxis);
    const axis = (s) => clamp(Math.round(s / spacing) + 1, 2, MAX_PROBES_PER_AXIS);
    return [axis(size[0]), axis(size[1]), axis(size[2])];
}

function boxOf(b) {
    if (!b) return null;
    if (b.min && b.max) {
        const lo = Array.isArray(b.min) ? b.min : [b.min.x, b.min.y, b.min.z];
        const hi = Array.isArray(b.max) ? b.max : [b.max.x, b.max.y, b.max.z];
        if (!(hi[0] >= lo[0] && hi[1] >= lo[1] && hi[2] >= lo[2])) return null;
        return { min: lo.slice(), max: hi.slice() };
    }
    return null;
}

function normalizeRes(r) {
    if (!r) return null;
    const rx = Number.isFinite(r.x) ? r.x : (Array.isArray(r) ? r[0] : NaN);
    const ry = Number.isFinite(r.y) ? r.y : (Array.isArray(r) ? r[1] : NaN);
    const rz = Number.isFinite(r.z) ? r.z : (Array.isArray(r) ? r[2] : NaN);
    if (!Number.isFinite(rx) || !Number.isFinite(ry) || !Number.isFinite(rz)) return null;
    const c = (v) => clamp(Math.round(v), 2, MAX_PROBES_PER_AXIS);
    return [c(rx), c(ry), c(rz)];
}

function createProbeField(opts) {
    opts = opts || {};
    const addon = opts.addon || loadAddon(opts.addonPath);
    let scene = opts.scene;
    if (!scene) throw new Error('createProbeField: { scene } is required.');

    let raysPerProbe = clamp(Math.round(Number(opts.rays || RAYS_PER_PROBE_DEFAULT) / 16) * 16 || RAYS_PER_PROBE_DEFAULT,
        RAYS_MIN, RAYS_MAX);
    let targetLongAxis = clamp(Math.round(Number(opts.divisions)) || TARGET_PROBES_LONG_AXIS, 2, MAX_PROBES_PER_AXIS);
    let cascades = (Math.round(Number(opts.cascades)) === 1) ? 1 : (opts.cascades == null ? 1 : 2);
    let jitterMode = opts.jitterMode === 'montecarlo' ? 'montecarlo' : 'gated';
    const jitterHysteresis = { gated: JITTER_HYSTERESIS_DEFAULTS.gated, montecarlo: JITTER_HYSTERESIS_DEFAULTS.montecarlo };
    let baseHysteresis = Number.isFinite(opts.hysteresis) ? clamp(opts.hysteresis, 0, 0.99) : jitterHysteresis[jitterMode];
    jitterHysteresis[jitterMode] = baseHysteresis;
    let hysteresisNormalize = false;
    let continuous = opts.continuous !== false;
    let detectSceneChanges = opts.autoDetectChanges !== false;
    let rayBudgetCeiling = RAYS_PER_TICK;
    let enabled = false;
    let manualVolumes = null;
    const blendPrecision = BLEND[opts.precision || 'bf16x3'];
    const traversal = TRAVERSAL[opts.traversal || 'ordered'];
    // clusteredLighting (js/gi_probes.js:1097): per-cell light lists for the NEE loop
    const clustered = opts.clusteredLighting === true;
    // reflectionQuality / roughReflections (resolveReflectionQuality, js/gi_probes.js:67-73)
    const reflectionName = opts.reflectionQuality == null
        ? (opts.roughReflections === true ? 'ultra' : 'off')
        : String(opts.reflectionQuality).trim().toLowerCase();
    if (!(reflectionName in REFLECTION)) {
        throw new RangeError(`Unknown reflectionQuality "${opts.reflectionQuality}". Expected off, rough, high, or ultra.`);
    }
    const reflections = REFLECTION[reflectionName];
    const glossyInterval = reflections === REFLECTION.high ? 2 : 1;
    const depthOct = opts.depthOct === 14 ? 14 : OCT_RES;
    const device = opts.device | 0;

    // live uniforms (the U set of js/gi_probes.js:1198-1255 that the update reads)
    const U = {
        filterStrength: 1, filterSmooth: 0.5, radianceClamp: 8, depthSharpness: 7, skyIntensity: 1,
        skySH: new Float32Array(27), nirGate: 0, nirGain: 1, classifyStrength: 0,
        tempChangeSigma1: 2.5, tempChangeHDrop: 0.30, tempClampSigma: 6, frameJitter: 0, hysteresisExponent: 1,
        emitterVisSeed: 0, skyConfigured: 0,
        reflectionSkyFallback: opts.reflectionSkyFallback === true ? 1 : 0,
    };
    // renderer-side knobs (state only)
    const R = {
        intensity: Number.isFinite(opts.intensity) ? opts.intensity : 1, reflectionIntensity: 1, roughnessLimit: 1,
        reflectionSkyFallback: false, chebyStrength: 1, normalDetail: 1, normalBias: 1,
    };
    let debugKnobs = {};

    // controller state
    let built = null;
    let field = null;
    let casc = [];
    let dirty = true;          // a rebuild is pending
    let buildRequestedFull = true;
    let inFlight = null;
    let frameCounter = 0;
    let emitterSeedCounter = 0;
    let solveIndex = 0;
    let disposed = false;
    let lastStats = null;
    let pendingAllTransforms = false;
    const pendingTransformTargets = new Set();
    let pendingDeformRefresh = false;
    let pendingLightRefresh = false;

    function requestRebuild(full) {
        dirty = true;
        if (full !== false) buildRequestedFull = true;
    }

    function placeGrid() {
        let lo, hi;
        const hasVolumes = Array.isArray(manualVolumes) && manualVolumes.length > 0;
        if (hasVolumes) {
            lo = [Infinity, Infinity, Infinity]; hi = [-Infinity, -Infinity, -Infinity];
            for (const v of manualVolumes) for (let k = 0; k < 3; k++) {
                lo[k] = Math.min(lo[k], v.box.min[k]); hi[k] = Math.max(hi[k], v.box.max[k]);
            }
        } else {
            const b = sceneBounds(built);
            lo = b.min; hi = b.max;
        }
        if (!(hi[0] >= lo[0])) return null;
        let gmin = lo.slice(), gsize = [hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]];
        if (!hasVolumes) {  // 6 % pad (js/gi_probes.js:3349-3352)
            const pad = gsize.map((s) => s * 0.06);
            gmin = gmin.map((g, k) => g - pad[k]);
            gsize = gsize.map((s, k) => s + 2 * pad[k]);
        }
        const resOverride = hasVolumes && manualVolumes.length === 1 && manualVolumes[0].res ? manualVolumes[0].res : null;
        const res = resOverride || computeGridResolution(gsize, targetLongAxis);
        const list = [{ gridMin: gmin, gridSize: gsize, res: res }];
        if (cascades > 1) {  // fine cascade: the central half of the box at twice the density
            const fmin = gmin.map((g, k) => g + gsize[k] * 0.25);
            const fsize = gsize.map((s) => s * 0.5);
            list.push({ gridMin: fmin, gridSize: fsize, res: res.map((r) => clamp(r, 2, MAX_PROBES_PER_AXIS)) });
        }
        return list;
    }

    function rebuild() {
        if (buildRequestedFull || !built) {
            built = buildScene(addon, scene);
        }
        const grids = placeGrid();
        if (!grids) throw new Error('createProbeField: the scene has no traceable bounds.');
        const same = field && casc.length === grids.length && casc.every((c, i) =>
            c.res.join() === grids[i].res.join() && c.gridMin.join() === grids[i].gridMin.join() &&
            c.gridSize.join() === grids[i].gridSize.join()) && field.rays === raysPerProbe;
        if (!same) {
            if (field) addon.fieldDestroy(field.h);
            const h = addon.fieldCreate({
                device: device, rays: raysPerProbe, depthOct: depthOct, precision: blendPrecision, traversal: traversal,
                flags: clustered ? FIELD_CLUSTERED : 0, reflections: reflections,
                cascades: grids.map((g) => ({
                    gridMin: Float32Array.from(g.gridMin), gridSize: Float32Array.from(g.gridSize),
                    res: Uint32Array.from(g.res),
                })),
            });
            field = { h: h, rays: raysPerProbe };
            casc = grids.map((g) => ({
                gridMin: g.gridMin, gridSize: g.gridSize, res: g.res, probeTotal: g.res[0] * g.res[1] * g.res[2],
                probeCursor: 0, needsClear: true, needsClassify: true, glossyPhase: 0,
            }));
        }
        addon.fieldSetScene(field.h, built.arrays);
        dirty = false;
        buildRequestedFull = false;
    }

    function updatedCap() {
        return Math.max(1, Math.min(MAX_PROBES_PER_TICK, Math.floor(rayBudgetCeiling / Math.max(1, raysPerProbe))));
    }

    // pending scene edits -> the library's ranged uploads and refits
    function applyPendingEdits() {
        if (!built || built.prebuilt) return;
        if (pendingAllTransforms || pendingTransformTargets.size) {
            const targets = pendingAllTransforms ? null : pendingTransformTargets;
            const slots = [], worlds = [], active = [];
            for (const inst of built.instances) {
                const obj = inst.object;
                const hit = !targets || targets.has(obj) || targets.has(inst.id) ||
                    (obj && obj.parent && targets.has(obj.parent));
                if (!hit) continue;
                const w = worldMatrix(obj);
                const slot = addon.sbSetTransform(built.builder, inst.id, w);
                slots.push(slot);
                for (let i = 0; i < 16; i++) worlds.push(w[i]);
                active.push(obj && obj.visible === false ? 0 : 1);
            }
            // { object, instanceIndex } entries address one instance directly
            if (targets) for (const t of targets) {
                if (t && t.object && Number.isInteger(t.instanceIndex)) {
                    const inst = built.instances[t.instanceIndex];
                    if (!inst) continue;
                    const w = worldMatrix(t.matrix ? t : inst.object);
                    slots.push(addon.sbSetTransform(built.builder, inst.id, w));
                    for (let i = 0; i < 16; i++) worlds.push(w[i]);
                    active.push(1);
                }
            }
            if (slots.length)
                addon.fieldSetTransforms(field.h, Uint32Array.from(slots), Float64Array.from(worlds), Uint8Array.from(active));
            pendingAllTransforms = false;
            pendingTransformTargets.clear();
        }
        if (pendingDeformRefresh) {
            // re-gather every mesh's current positions into the pooled vertex data, one
            // ranged upload per mesh, then refit every BLAS range and the TLAS
            const vd = built.arrays.vertexData;
            const ranges = [];
            for (const m of built.meshes) {
                const pos = m.source.geometry ? m.source.geometry.attributes.position.array : m.source.positions;
                for (let v = 0; v < m.vcount; v++) for (let k = 0; k < 3; k++) vd[(m.vertBase + v) * 8 + k] = pos[v * 3 + k];
                ranges.push(m.vertBase * 8, m.vcount * 8);
            }
            if (ranges.length) addon.fieldUpdateRanges(field.h, BUF.VERTEX_DATA, Uint32Array.from(ranges), vd);
            const a = built.arrays, roots = new Set(), br = [];
            for (let i = 0; i < a.instCount; i++) {
                const o = a.instBase + i * 28;
                const key = a.materials[o + 12] + ':' + a.materials[o + 13];
                if (roots.has(key)) continue;
                roots.add(key);
                br.push(a.materials[o + 12], a.materials[o + 13]);
            }
            addon.fieldRefit(field.h, Uint32Array.from(br), true);
            pendingDeformRefresh = false;
        }
    }

    function tickParams(C, ci, updated) {
        const p = {
            cascade: ci, probeOffset: C.probeCursor, updatedCount: updated, needsClear: C.needsClear ? 1 : 0,
            frameJitter: U.frameJitter, hysteresis: baseHysteresis, hysteresisExponent: U.hysteresisExponent,
            depthSharpness: U.depthSharpness, radianceClamp: U.radianceClamp, classifyStrength: U.classifyStrength,
            filterStrength: U.filterStrength, filterSmooth: U.filterSmooth, tempChangeSigma1: U.tempChangeSigma1,
            tempChangeHDrop: U.tempChangeHDrop, tempClampSigma: U.tempClampSigma, nirGate: U.nirGate, nirGain: U.nirGain,
            skyIntensity: U.skyIntensity, skySH: U.skySH, emitterVisSeed: U.emitterVisSeed,
            skyConfigured: U.skyConfigured, reflectionSkyFallback: U.reflectionSkyFallback,
            glossyPhase: C.glossyPhase || 0,
        };
        for (const k of Object.keys(debugKnobs)) p[k] = debugKnobs[k];
        return p;
    }

    function waitTicket(ticket) {
        return new Promise(function (resolve, reject) {
            function poll() {
                if (disposed) { resolve(null); return; }
                let done;
                try { done = addon.fieldPoll(field.h, ticket); } catch (e) { reject(e); return; }
                if (done) resolve(ticket); else setImmediate(poll);
            }
            poll();
        });
    }

    // tick (js/gi_probes.js:3472-3849): one bounded solve of the next probe range of one
    // cascade per call; rebuilds and scene edits are applied first.
    function tick(state) {
        if (disposed || !enabled) return Promise.resolve(null);
        if (inFlight) return inFlight;
        try {
            if (dirty || !field) rebuild();
            applyPendingEdits();
            if (jitterMode === 'montecarlo') U.frameJitter = fract(frameCounter * GOLDEN_RATIO_FRACT);
            frameCounter++;
            const ci = casc.length > 1 ? solveIndex++ % casc.length : 0;
            // the emitter sample target moves with the Monte Carlo epoch of cascade 0
            // (js/gi_probes.js:3767-3777)
            if (jitterMode === 'montecarlo' && ci === 0) {
                emitterSeedCounter = (emitterSeedCounter + 1) >>> 0;
                U.emitterVisSeed = fract(emitterSeedCounter * GOLDEN_RATIO_FRACT);
            }
            const C = casc[ci];
            const updated = Math.min(updatedCap(), C.probeTotal);
            // classifyKernel runs once per rebuild while classifyStrength > 0
            // (js/gi_probes.js:3803-3806)
            if (C.needsClassify) {
                if (U.classifyStrength > 0) addon.fieldClassify(field.h, ci);
                C.needsClassify = false;
            }
            const ticket = addon.fieldTick(field.h, tickParams(C, ci, updated));
            C.glossyPhase = ((C.glossyPhase || 0) + 1) % glossyInterval;  // js/gi_probes.js:3831-3835
            C.needsClear = false;
            C.probeCursor = (C.probeCursor + updated) % C.probeTotal;
            inFlight = waitTicket(ticket).then(function (t) {
                inFlight = null;
                if (t !== null) lastStats = addon.fieldStats(field.h);
                return { cascade: ci, updated: updated, stats: lastStats };
            }, function (e) {
                inFlight = null;
                if (e && (e.code === STATUS.E_LAYOUT || e.code === STATUS.E_CAPACITY)) requestRebuild(true);
                throw e;
            });
            return inFlight;
        } catch (e) {
            if (e && (e.code === STATUS.E_LAYOUT || e.code === STATUS.E_CAPACITY)) requestRebuild(true);
            return Promise.reject(e);
        }
    }

    function markTransformsDirty(targets) {
        if (targets == null) { pendingAllTransforms = true; pendingTransformTargets.clear(); return; }
        if (pendingAllTransforms) return;
        const list = Array.isArray(targets) || targets instanceof Set ? targets : [targets];
        for (const t of list) if (t != null) pendingTransformTargets.add(t);
    }
    function markDeformsDirty() { pendingDeformRefresh = true; }
    function markMaterialValuesDirty() { requestRebuild(true); }
    function markTopologyDirty() { requestRebuild(true); }
    function forceLightingRefresh() { pendingLightRefresh = true; requestRebuild(true); }
    function notifySceneChange(change) {
        if (Array.isArray(change)) { for (const c of change) notifySceneChange(c); return true; }
        const type = typeof change === 'string' ? change : (change && change.type);
        const pick = (c) => c == null ? null : (c.objects != null ? c.objects : c.object != null ? c.object :
            c.target != null ? c.target : null);
        switch (String(type || '').toLowerCase()) {
            case 'transform': case 'xform': markTransformsDirty(pick(change)); return true;
            case 'deform': case 'vertices': markDeformsDirty(); return true;
            case 'light': case 'lighting': forceLightingRefresh(); return true;
            case 'material': case 'materials': markMaterialValuesDirty(); return true;
            case 'add': case 'added': case 'remove': case 'removed': case 'topology': case 'geometry':
            case 'structure': markTopologyDirty(); return true;
            default: return false;
        }
    }
    function setVolumes(boxes) {
        const arr = Array.isArray(boxes) ? boxes : [boxes];
        const list = [];
        for (const entry of arr) {
            if (!entry) continue;
            const box = boxOf(entry.box || entry);
            if (!box) continue;
            list.push({ box: box, res: entry.box ? normalizeRes(entry.res) : null });
        }
        manualVolumes = list.length ? list : null;
        requestRebuild(false);
    }

    const debugReadIds = { irr: BUF.IRR_STATE, depth: BUF.DEPTH_STATE, state: BUF.META, rays: BUF.RAY_DATA,
        mat: BUF.MATERIALS, lights: BUF.LIGHTS };

    const api = {
        node: null,
        tick: tick,
        resetFramePacing: function () {},
        setEnabled: function (on) { enabled = on !== false; },
        getEnabled: function () { return enabled; },
        setIntensity: function (v) { if (Number.isFinite(v)) R.intensity = v; },
        getIntensity: function () { return R.intensity; },
        setReflectionIntensity: function (v) { if (Number.isFinite(v)) R.reflectionIntensity = v; },
        getReflectionIntensity: function () { return R.reflectionIntensity; },
        setRoughnessLimit: function (v) { if (Number.isFinite(v)) R.roughnessLimit = v; },
        getRoughnessLimit: function () { return R.roughnessLimit; },
        getReflectionQuality: function () { return reflectionName; },
        setChebyStrength: function (v) { if (Number.isFinite(v)) R.chebyStrength = v; },
        setNormalDetail: function (v) { if (Number.isFinite(v)) R.normalDetail = v; },
        setClassifyStrength: function (v) {
            if (!Number.isFinite(v)) return;
            U.classifyStrength = clamp(v, 0, 1);
            // the state buffer is only written while the strength is positive (js/gi_probes.js:4570-4577)
            if (U.classifyStrength > 0) for (const C of casc) C.needsClassify = true;
        },
        setDivisions: function (n) {
            const v = clamp(Math.round(Number(n)) || TARGET_PROBES_LONG_AXIS, 2, MAX_PROBES_PER_AXIS);
            if (v === targetLongAxis) return;
            targetLongAxis = v;
            requestRebuild(false);
        },
        getDivisions: function () { return targetLongAxis; },
        setRays: function (n) {
            const v = clamp(Math.round(Number(n) / 16) * 16 || RAYS_PER_PROBE_DEFAULT, RAYS_MIN, RAYS_MAX);
            if (v === raysPerProbe) return;
            raysPerProbe = v;
            requestRebuild(false);
        },
        getRays: function () { return raysPerProbe; },
        setRayBudget: function (v) {
            if (!Number.isFinite(v)) return;
            const budget = clamp(Math.round(v), RAYS_PER_TICK_MIN, 524288);
            if (budget === rayBudgetCeiling) return;
            rayBudgetCeiling = budget;
        },
        getRayBudget: function () { return rayBudgetCeiling; },
        setFilterStrength: function (v) { if (Number.isFinite(v)) U.filterStrength = clamp(v, 0, 1); },
        setSmoothness: function (v) { if (Number.isFinite(v)) U.filterSmooth = clamp(v, 0, 1); },
        setHysteresis: function (v) {
            if (!Number.isFinite(v)) return;
            baseHysteresis = clamp(v, 0, 0.99);
            jitterHysteresis[jitterMode] = baseHysteresis;
        },
        getHysteresis: function () { return baseHysteresis; },
        setHysteresisNormalization: function (on) {
            hysteresisNormalize = on !== false;
            if (!hysteresisNormalize) U.hysteresisExponent = 1;
        },
        getHysteresisNormalization: function () { return hysteresisNormalize; },
        setJitterMode: function (mode) {
            const m = String(mode == null ? '' : mode).toLowerCase().replace(/[\s_-]/g, '');
            const next = (m === 'gated' || m === 'gatedbasis') ? 'gated' : (m === 'montecarlo' || m === 'mc') ? 'montecarlo' : null;
            if (!next || next === jitterMode) return;
            jitterHysteresis[jitterMode] = baseHysteresis;
            jitterMode = next;
            baseHysteresis = jitterHysteresis[jitterMode];
            if (jitterMode === 'gated') U.frameJitter = 0;
        },
        getJitterMode: function () { return jitterMode; },
        setNormalBias: function (v) { if (Number.isFinite(v)) R.normalBias = clamp(v, 0, 8); },
        setRadianceClamp: function (v) { if (Number.isFinite(v)) U.radianceClamp = Math.max(0, v); },
        setDepthSharpness: function (v) { if (Number.isFinite(v)) U.depthSharpness = Math.max(0.01, clamp(v, 0, 200)); },
        // null | 27 SH-9 radiance coefficients (three's SphericalHarmonics3 order, rgb per
        // coefficient) | { coefficients: [9 x {x,y,z}] } | a uniform colour
        setSky: function (sky) {
            const sh = new Float32Array(27);
            U.skyConfigured = sky == null ? 0 : 1;  // js/gi_probes.js:1316
            if (sky == null) { U.skySH = sh; return; }
            if (ArrayBuffer.isView(sky) || Array.isArray(sky)) {
                if (sky.length === 27) { for (let i = 0; i < 27; i++) sh[i] = sky[i]; U.skySH = sh; return; }
            }
            const co = sky.coefficients || (sky.sh && sky.sh.coefficients);
            if (co && co.length === 9) {
                for (let i = 0; i < 9; i++) { sh[3 * i] = co[i].x; sh[3 * i + 1] = co[i].y; sh[3 * i + 2] = co[i].z; }
                U.skySH = sh; return;
            }
            const c = (typeof sky === 'number' || (sky && typeof sky.r === 'number')) ? sky : null;
            if (c != null) {  // uniform radiance L: c0 = L / Y00
                const col = typeof c === 'number' ? [((c >> 16) & 255) / 255, ((c >> 8) & 255) / 255, (c & 255) / 255] : [c.r, c.g, c.b];
                for (let a = 0; a < 3; a++) sh[a] = col[a] / 0.282095;
                U.skySH = sh;
            }
        },
        setSkyIntensity: function (v) { if (Number.isFinite(v)) U.skyIntensity = Math.max(0, v); },
        setReflectionSkyFallback: function (on) { U.reflectionSkyFallback = on === true ? 1 : 0; },
        getReflectionSkyFallback: function () { return U.reflectionSkyFallback > 0.5; },
        hasRoughReflections: function () { return reflections !== REFLECTION.off; },
        hasGlossyReflections: function () { return reflections === REFLECTION.high || reflections === REFLECTION.ultra; },
        getSkyIntensity: function () { return U.skyIntensity; },
        setNirSensing: function (on) { U.nirGate = on ? 1 : 0; },
        getNirSensing: function () { return U.nirGate > 0.5; },
        setNirGain: function (gain) { U.nirGain = Number.isFinite(gain) ? Math.max(0, gain) : 1; },
        getNirGain: function () { return U.nirGain; },
        setChangeThreshold: function (v) { if (Number.isFinite(v)) U.tempChangeSigma1 = clamp(v, 0.5, 8); },
        setSnapAmount: function (v) { if (Number.isFinite(v)) U.tempChangeHDrop = clamp(v, 0, 0.9); },
        setFireflyClamp: function (v) { if (Number.isFinite(v)) U.tempClampSigma = clamp(v, 1, 20); },
        getChangeThreshold: function () { return U.tempChangeSigma1; },
        getSnapAmount: function () { return U.tempChangeHDrop; },
        getFireflyClamp: function () { return U.tempClampSigma; },
        setDebugKnobs: function (k) { debugKnobs = Object.assign({}, debugKnobs, k || {}); },
        getDebugKnobs: function () { return Object.assign({}, debugKnobs); },
        forceLightingRefresh: forceLightingRefresh,
        setCascades: function (n) {
            const v = (Math.round(Number(n)) === 1) ? 1 : 2;
            if (v === cascades) return;
            cascades = v;
            requestRebuild(false);
        },
        getCascades: function () { return cascades; },
        setContinuous: function (on) { continuous = on === true; },
        getContinuous: function () { return continuous; },
        setAutoDetectChanges: function (on) { detectSceneChanges = on !== false; },
        getAutoDetectChanges: function () { return detectSceneChanges; },
        requestRebuild: requestRebuild,
        markTransformsDirty: markTransformsDirty,
        markDeformsDirty: markDeformsDirty,
        markMaterialValuesDirty: markMaterialValuesDirty,
        markTopologyDirty: markTopologyDirty,
        notifySceneChange: notifySceneChange,
        setBounds: function (box) { setVolumes(box ? [box] : null); },
        setVolumes: setVolumes,
        isSupported: function () { return true; },
        hasData: function () { return lastStats !== null; },
        getStats: function (ci) {
            const C = casc[ci || 0] || casc[0];
            if (!C) return null;
            return {
                probes: C.probeTotal, res: C.res.slice(), atlas: [C.res[0] * TILE, C.res[1] * C.res[2] * TILE],
                rays: raysPerProbe, oct: OCT_RES, tile: TILE, active: enabled, cascades: cascades, cascade: ci || 0,
                budgetRays: rayBudgetCeiling, device: lastStats,
            };
        },
        getResolution: function (ci) { const C = casc[ci || 0] || casc[0]; return C ? C.res.slice() : null; },
        getBounds: function (ci) {
            const C = casc[ci || 0] || casc[0];
            if (!C) return null;
            return { min: C.gridMin.slice(), max: C.gridMin.map((g, k) => g + C.gridSize[k]) };
        },
        _debugRead: function (which, ci) {
            const id = debugReadIds[which];
            if (id == null || !field) return Promise.resolve(null);
            try { return Promise.resolve(new Float32Array(addon.fieldRead(field.h, id, ci || 0))); }
            catch (e) { return Promise.resolve({ error: String(e) }); }
        },
        _debugState: function () {
            return {
                dirty: dirty, inFlight: !!inFlight, baseHysteresis: baseHysteresis, jitterMode: jitterMode,
                frameJitter: U.frameJitter, frameCounter: frameCounter, rays: raysPerProbe, cascades: cascades,
                pendingTransformCount: pendingAllTransforms ? -1 : pendingTransformTargets.size,
                pendingDeformRefresh: pendingDeformRefresh, pendingLightRefresh: pendingLightRefresh,
                cascadeState: casc.map((C) => ({ probes: C.probeTotal, cursor: C.probeCursor, needsClear: C.needsClear })),
            };
        },
        _debugTickParams: function (ci) {
            const C = casc[ci || 0];
            return C ? tickParams(C, ci || 0, Math.min(updatedCap(), C.probeTotal)) : null;
        },
        dispose: function () {
            if (disposed) return;
            disposed = true;
            if (field) addon.fieldDestroy(field.h);
            if (built && built.builder) addon.sbDestroy(built.builder);
            field = null;
        },
    };
    return api;
}

module.exports = { createProbeField, computeGridResolution };
