'use strict';
// Host-side tests of the JS shim (no GPU): argument normalisation, change routing,
// scene ingestion through the real addon's scene builder, and the calls tick() makes.
// The field-level addon calls (which need a GPU) are replaced by a recorder.
const assert = require('assert');
const path = require('path');
const shim = require(path.join(__dirname, '..', '..', 'speedball_amd', 'js'));

const real = shim.loadAddon();

function recorder() {
    const calls = [];
    let ticket = 0;
    const fake = Object.assign({}, real, {
        fieldCreate: (d) => { calls.push(['create', d]); return { fake: true }; },
        fieldDestroy: () => calls.push(['destroy']),
        fieldSetScene: (h, s) => calls.push(['setScene', s]),
        fieldTick: (h, p) => { calls.push(['tick', p]); return ++ticket; },
        fieldPoll: () => true,
        fieldStats: () => ({ msTotal: 1 }),
        fieldSetTransforms: (h, slots, worlds, active) => calls.push(['transforms', slots, worlds, active]),
        fieldUpdateRanges: (h, id, ranges, src) => calls.push(['ranges', id, ranges, src.length]),
        fieldRefit: (h, ranges, tl) => calls.push(['refit', ranges, tl]),
        fieldClassify: (h, c) => calls.push(['classify', c]),
        fieldRead: () => new ArrayBuffer(16),
    });
    return { fake, calls };
}

function boxMesh(sx, sy, sz) {
    const p = [];
    for (let i = 0; i < 8; i++) p.push((i & 1 ? 0.5 : -0.5) * sx, (i & 2 ? 0.5 : -0.5) * sy, (i & 4 ? 0.5 : -0.5) * sz);
    const f = [[0, 2, 3, 1], [4, 5, 7, 6], [0, 1, 5, 4], [2, 6, 7, 3], [0, 4, 6, 2], [1, 3, 7, 5]];
    const idx = [];
    for (const q of f) idx.push(q[0], q[1], q[2], q[0], q[2], q[3]);
    return { positions: new Float32Array(p), indices: new Uint32Array(idx) };
}

function translation(x, y, z) {
    const m = new Float64Array(16);
    m[0] = m[5] = m[10] = m[15] = 1;
    m[12] = x; m[13] = y; m[14] = z;
    return m;
}

function plainScene() {
    const a = boxMesh(1, 1, 1), b = boxMesh(2, 0.5, 1);
    return {
        materials: [{ color: [0.5, 0.5, 0.5] }, { color: { r: 0.8, g: 0.2, b: 0.1 }, side: 2 }],
        meshes: [Object.assign(a, { materialIndex: 0 }), Object.assign(b, { materialIndex: 1 })],
        instances: [{ mesh: 0, matrix: translation(0, 0, 0) }, { mesh: 1, matrix: translation(3, 1, 0) },
            { mesh: 0, matrix: translation(-2, 0, 2) }],
        lights: [{ type: 'point', position: [0, 3, 0], color: [1, 1, 1], intensity: 10 }],
    };
}

// a three-like graph of the same scene
function threeScene() {
    const s = plainScene();
    const geoms = s.meshes.map((m) => ({ attributes: { position: { array: m.positions, version: 0 } }, index: { array: m.indices } }));
    const objs = s.instances.map((inst) => ({
        isMesh: true, geometry: geoms[inst.mesh], material: s.materials[inst.mesh === 0 ? 0 : 1],
        matrixWorld: { elements: Array.from(inst.matrix) },
    }));
    const light = { isPointLight: true, matrixWorld: { elements: Array.from(translation(0, 3, 0)) }, color: { r: 1, g: 1, b: 1 }, intensity: 10 };
    const all = objs.concat([light]);
    return { traverse: (cb) => all.forEach(cb), objs: objs };
}

const tests = {};

tests.setters_normalise_like_the_reference = () => {
    const { fake } = recorder();
    const gi = shim.createProbeField({ scene: plainScene(), addon: fake });
    gi.setRays(70); assert.strictEqual(gi.getRays(), 64);
    gi.setRays(1000); assert.strictEqual(gi.getRays(), 256);
    gi.setRays(20); assert.strictEqual(gi.getRays(), 32);
    gi.setRays(0); assert.strictEqual(gi.getRays(), 64);      // Math.round(0/16)*16 || default
    gi.setRays(200); assert.strictEqual(gi.getRays(), 208);   // rounds to 16
    gi.setHysteresis(1.5); assert.strictEqual(gi.getHysteresis(), 0.99);
    gi.setHysteresis(-1); assert.strictEqual(gi.getHysteresis(), 0);
    gi.setHysteresis(NaN); assert.strictEqual(gi.getHysteresis(), 0);
    gi.setDivisions(100); assert.strictEqual(gi.getDivisions(), 32);
    gi.setDivisions('abc'); assert.strictEqual(gi.getDivisions(), 12);
    gi.setDivisions(1); assert.strictEqual(gi.getDivisions(), 2);
    gi.setRayBudget(10); assert.strictEqual(gi.getRayBudget(), 2048);
    gi.setRayBudget(1e9); assert.strictEqual(gi.getRayBudget(), 524288);
    gi.setCascades(1); assert.strictEqual(gi.getCascades(), 1);
    gi.setCascades(7); assert.strictEqual(gi.getCascades(), 2);
    gi.setJitterMode('Monte-Carlo'); assert.strictEqual(gi.getJitterMode(), 'montecarlo');
    assert.strictEqual(gi.getHysteresis(), 0.90);            // montecarlo's remembered profile
    gi.setJitterMode('gated basis'); assert.strictEqual(gi.getJitterMode(), 'gated');
    assert.strictEqual(gi.getHysteresis(), 0);               // gated remembers the override above
    gi.setChangeThreshold(100); assert.strictEqual(gi.getChangeThreshold(), 8);
    gi.setSnapAmount(-1); assert.strictEqual(gi.getSnapAmount(), 0);
    gi.setFireflyClamp(0); assert.strictEqual(gi.getFireflyClamp(), 1);
    gi.setNirGain(-3); assert.strictEqual(gi.getNirGain(), 0);
    gi.setNirGain('x'); assert.strictEqual(gi.getNirGain(), 1);
    gi.setSkyIntensity(-2); assert.strictEqual(gi.getSkyIntensity(), 0);
};

tests.scene_change_routing = () => {
    const { fake } = recorder();
    const gi = shim.createProbeField({ scene: plainScene(), addon: fake });
    assert.strictEqual(gi.notifySceneChange('transform'), true);
    assert.strictEqual(gi._debugState().pendingTransformCount, -1);
    assert.strictEqual(gi.notifySceneChange({ type: 'deform' }), true);
    assert.strictEqual(gi._debugState().pendingDeformRefresh, true);
    assert.strictEqual(gi.notifySceneChange('light'), true);
    assert.strictEqual(gi.notifySceneChange('bogus'), false);
    assert.strictEqual(gi.notifySceneChange(['add', 'material']), true);
    const gi2 = shim.createProbeField({ scene: plainScene(), addon: fake });
    const o = {};
    gi2.markTransformsDirty(o);
    gi2.markTransformsDirty([{}, {}]);
    assert.strictEqual(gi2._debugState().pendingTransformCount, 3);
    gi2.markTransformsDirty(null);
    assert.strictEqual(gi2._debugState().pendingTransformCount, -1);
};

tests.plain_and_three_like_scenes_build_the_same_buffers = () => {
    const a = shim.buildScene(real, plainScene()).arrays;
    const b = shim.buildScene(real, threeScene()).arrays;
    for (const k of ['bvhNodes', 'triIndex', 'vertexData', 'triMaterial', 'materials', 'lights']) {
        assert.deepStrictEqual(Array.from(a[k]), Array.from(b[k]), k);
    }
    assert.strictEqual(a.instCount, 3);
    assert.strictEqual(a.lights.length, 17);
    assert.strictEqual(a.lights[0], 1);            // point light
    assert.strictEqual(a.lights[7], 10);           // colour x intensity
};

tests.tick_calls_and_grid_placement = async () => {
    const { fake, calls } = recorder();
    const gi = shim.createProbeField({ scene: plainScene(), addon: fake, rays: 64, divisions: 12 });
    gi.setEnabled(true);
    const r = await gi.tick({});
    const create = calls.find((c) => c[0] === 'create')[1];
    // computeGridResolution of the padded TLAS-root bounds (js/gi_probes.js:938-943, 3349-3352)
    const built = shim.buildScene(real, plainScene()).arrays;
    const t = built.materials, base = built.tlasBase;
    const size = [0, 1, 2].map((k) => (t[base + 3 + k] - t[base + k]) * 1.12);
    const res = shim.computeGridResolution(size, 12);
    assert.deepStrictEqual(Array.from(create.cascades[0].res), res);
    const tick = calls.filter((c) => c[0] === 'tick');
    assert.strictEqual(tick.length, 1);
    const p = tick[0][1];
    assert.strictEqual(p.needsClear, 1);
    assert.strictEqual(p.probeOffset, 0);
    assert.strictEqual(p.updatedCount, Math.min(1536, res[0] * res[1] * res[2]));
    assert.strictEqual(p.hysteresis, 0.60);
    assert.strictEqual(p.frameJitter, 0);
    assert.strictEqual(r.updated, p.updatedCount);
    await gi.tick({});
    const p2 = calls.filter((c) => c[0] === 'tick')[1][1];
    assert.strictEqual(p2.needsClear, 0);
    // in-flight guard: two calls before completion share one promise
    fake.fieldPoll = () => false;
    const a = gi.tick({}), b = gi.tick({});
    assert.strictEqual(a, b);
    fake.fieldPoll = () => true;
    await a;
};

tests.transform_and_deform_edits_reach_the_library = async () => {
    const { fake, calls } = recorder();
    const s = threeScene();
    const gi = shim.createProbeField({ scene: s, addon: fake });
    gi.setEnabled(true);
    await gi.tick({});
    s.objs[1].matrixWorld.elements[12] = 5;        // move the second instance
    gi.markTransformsDirty(s.objs[1]);
    await gi.tick({});
    const tr = calls.filter((c) => c[0] === 'transforms');
    assert.strictEqual(tr.length, 1);
    assert.strictEqual(tr[0][1].length, 1);
    assert.strictEqual(tr[0][2][12], 5);
    s.objs[0].geometry.attributes.position.array[0] = -0.75;
    gi.markDeformsDirty();
    await gi.tick({});
    const rg = calls.filter((c) => c[0] === 'ranges');
    assert.strictEqual(rg.length, 1);
    assert.strictEqual(rg[0][1], shim.BUF.VERTEX_DATA);
    const rf = calls.filter((c) => c[0] === 'refit');
    assert.strictEqual(rf.length, 1);
    assert.strictEqual(rf[0][1].length, 4);        // two unique BLAS ranges
};

tests.classification_runs_once_per_rebuild_while_the_strength_is_positive = async () => {
    // js/gi_probes.js:3803-3806 (run when a cascade needs it and classifyStrength > 0),
    // :4570-4577 (a positive strength asks every cascade again)
    const { fake, calls } = recorder();
    const gi = shim.createProbeField({ scene: plainScene(), addon: fake, cascades: 1 });
    gi.setEnabled(true);
    await gi.tick({});
    await gi.tick({});
    assert.strictEqual(calls.filter((c) => c[0] === 'classify').length, 0);  // strength 0: never
    gi.setClassifyStrength(0.5);
    await gi.tick({});
    await gi.tick({});
    const cl = calls.filter((c) => c[0] === 'classify');
    assert.strictEqual(cl.length, 1);
    assert.strictEqual(cl[0][1], 0);
    const order = calls.map((c) => c[0]);
    assert.ok(order.lastIndexOf('classify') < order.lastIndexOf('tick'));
    gi.setClassifyStrength(0.75);
    await gi.tick({});
    assert.strictEqual(calls.filter((c) => c[0] === 'classify').length, 2);
    gi.setClassifyStrength(0);
    await gi.tick({});
    assert.strictEqual(calls.filter((c) => c[0] === 'classify').length, 2);
};

tests.layout_errors_request_a_rebuild = async () => {
    const { fake, calls } = recorder();
    const gi = shim.createProbeField({ scene: plainScene(), addon: fake });
    gi.setEnabled(true);
    await gi.tick({});
    fake.fieldTick = () => { const e = new Error('sbgi error -2: layout'); e.code = -2; throw e; };
    let failed = false;
    await gi.tick({}).catch((e) => { failed = e.code === -2; });
    assert.ok(failed);
    assert.strictEqual(gi._debugState().dirty, true);
    fake.fieldTick = (h, p) => { calls.push(['tick', p]); return 99; };
    await gi.tick({});
    assert.strictEqual(calls.filter((c) => c[0] === 'setScene').length, 2);
};

tests.emitters_and_clustered_lighting = async () => {
    // a giEmitter mesh of a three-like graph and its plain { center, radius, area, radiance }
    // description give the same type-3 record, placed after the analytic lights
    const s = threeScene();
    const panel = s.objs[1];                       // the 2 x 0.5 x 1 box at (3, 1, 0)
    panel.userData = { giEmitter: true };
    panel.material = Object.assign({}, panel.material, { emissive: [1, 0.5, 0.25], emissiveIntensity: 4 });
    const three = shim.buildScene(real, s).arrays;
    assert.strictEqual(three.emitterCount, 1);
    assert.strictEqual(three.emitterBase, 1);
    const rec = three.lights.subarray(17, 34);
    const area = 2 * (2 * 0.5 + 2 * 1 + 0.5 * 1), radius = Math.hypot(1, 0.25, 0.5);
    const near = (a, b) => Math.abs(a - b) <= 1e-5 * Math.max(1, Math.abs(b));
    assert.strictEqual(rec[0], 3);
    assert.ok(near(rec[1], 3) && near(rec[2], 1) && near(rec[3], 0));
    assert.ok(near(rec[7], 4) && near(rec[8], 2) && near(rec[9], 1));
    assert.ok(near(rec[14], radius) && near(rec[15], area / 4));
    assert.ok(near(rec[10], Math.sqrt(4 * (area / 4) / 0.005)));
    const plain = plainScene();
    plain.emitters = [{ center: [3, 1, 0], radius: radius, area: area, radiance: [4, 2, 1] }];
    const p = shim.buildScene(real, plain).arrays;
    assert.deepStrictEqual(Array.from(p.lights), Array.from(three.lights));
    // clusteredLighting sets the create flag; Monte Carlo ticks of cascade 0 move the
    // emitter sample seed, gated ticks hold it
    const { fake, calls } = recorder();
    const gi = shim.createProbeField({ scene: plain, addon: fake, clusteredLighting: true, cascades: 1 });
    gi.setEnabled(true);
    await gi.tick({});
    assert.strictEqual(calls.find((c) => c[0] === 'create')[1].flags, 2);
    assert.strictEqual(calls.filter((c) => c[0] === 'tick')[0][1].emitterVisSeed, 0);
    gi.setJitterMode('montecarlo');
    await gi.tick({});
    await gi.tick({});
    const seeds = calls.filter((c) => c[0] === 'tick').map((c) => c[1].emitterVisSeed);
    assert.ok(seeds[1] > 0 && seeds[2] > 0 && seeds[1] !== seeds[2]);
};

tests.reflection_quality_and_glossy_phase = async () => {
    // reflectionQuality resolves like the reference; HIGH alternates the glossy phase per
    // solve of a cascade; setSky marks the sky configured for the reflection fallback
    assert.throws(() => shim.createProbeField({ scene: plainScene(), addon: recorder().fake, reflectionQuality: 'max' }),
        RangeError);
    const { fake, calls } = recorder();
    const gi = shim.createProbeField({ scene: plainScene(), addon: fake, reflectionQuality: 'High', cascades: 1,
        reflectionSkyFallback: true });
    assert.strictEqual(gi.getReflectionQuality(), 'high');
    assert.ok(gi.hasRoughReflections() && gi.hasGlossyReflections() && gi.getReflectionSkyFallback());
    gi.setEnabled(true);
    await gi.tick({});
    gi.setSky([1, 1, 1]);
    await gi.tick({});
    await gi.tick({});
    assert.strictEqual(calls.find((c) => c[0] === 'create')[1].reflections, 2);
    const ticks = calls.filter((c) => c[0] === 'tick').map((c) => c[1]);
    assert.deepStrictEqual(ticks.map((p) => p.glossyPhase), [0, 1, 0]);
    assert.deepStrictEqual(ticks.map((p) => p.skyConfigured), [0, 1, 1]);
    assert.ok(ticks.every((p) => p.reflectionSkyFallback === 1));
    const off = shim.createProbeField({ scene: plainScene(), addon: recorder().fake, roughReflections: false });
    assert.ok(!off.hasRoughReflections());
    const ultra = shim.createProbeField({ scene: plainScene(), addon: recorder().fake, roughReflections: true });
    assert.strictEqual(ultra.getReflectionQuality(), 'ultra');
};

tests.install_handle_forwards_and_guards = async () => {
    const { fake } = recorder();
    let errs = 0;
    const h = shim.installSpeedballGI({ scene: plainScene(), addon: fake, onError: () => { errs++; } });
    h.setRays(100);
    assert.strictEqual(h.getRays(), 96);
    assert.strictEqual(typeof h.markTransformsDirty, 'function');
    assert.strictEqual(typeof h.advanced.debug.read, 'function');
    await h.update();
    fake.fieldTick = () => { throw new Error('boom'); };
    await h.update(); await h.update();
    assert.strictEqual(errs, 1);                   // once per failure streak
    h.dispose();
    assert.strictEqual(h.update(), undefined);
    assert.throws(() => shim.installSpeedballGI({}), /scene/);
};

tests.addon_rejects_malformed_scene_arrays = () => {
    // the real addon's fieldSetScene checks the scene's element types and lengths before it
    // looks at the field, so a malformed scene is refused without a device
    const valid = () => ({
        bvhNodes: new Uint32Array(8), triIndex: new Uint32Array(3), vertexData: new Float32Array(24),
        triMaterial: new Uint32Array(1), materials: new Float32Array(28 + 28 + 12), lights: new Float32Array(17),
        instBase: 28, instCount: 1, tlasBase: 56, tlasNodeCount: 1,
    });
    assert.throws(() => real.fieldSetScene(null, valid()), /bad field/);
    const bad = [
        ['bvhNodes', (s) => { s.bvhNodes = new Uint8Array(32); }],
        ['bvhNodes', (s) => { s.bvhNodes = new Uint32Array(12); }],
        ['triIndex', (s) => { s.triIndex = new Uint16Array(3); }],
        ['vertexData', (s) => { s.vertexData = new Uint32Array(24); }],
        ['triMaterial', (s) => { s.triMaterial = new Uint32Array(2); }],
        ['triMaterial', (s) => { s.triMaterial = new Uint8Array(1); }],
        ['materials', (s) => { s.instCount = 2; }],
        ['materials', (s) => { s.tlasNodeCount = 2; }],
        ['materials', (s) => { s.materials = new Float64Array(68); }],
        ['lights', (s) => { s.lights = new Float32Array(16); }],
        ['lights', (s) => { s.emitterBase = 0; s.emitterCount = 2; }],
        ['triIndex', (s) => { delete s.triIndex; }],
    ];
    for (const [name, edit] of bad) {
        const s = valid();
        edit(s);
        assert.throws(() => real.fieldSetScene(null, s), new RegExp('malformed scene: ' + name));
    }
    const noLights = valid();
    delete noLights.lights;
    assert.throws(() => real.fieldSetScene(null, noLights), /bad field/);   // lights are optional
};

(async () => {
    const only = process.argv[2];
    let n = 0;
    for (const name of Object.keys(tests)) {
        if (only && name !== only) continue;
        await tests[name]();
        console.log('ok', name);
        n++;
    }
    console.log('passed', n);
})().catch((e) => { console.error(e && e.stack || e); process.exit(1); });
