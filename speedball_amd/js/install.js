'use strict';
// installSpeedballGI (js/install.js:216-504): one-call setup returning the handle whose
// update() the host calls once per frame. There is no WebGPU renderer here: `renderer`
// is accepted and ignored, and the lights-node factory (renderer-side) is not installed.
// The handle exposes the same names, forwarding to createProbeField's api, with the
// same disposed-guard, onError-once-per-streak and camera idle tracking.
const { createProbeField } = require('./probe_field');

function now() {
    const t = process.hrtime();
    return t[0] * 1000 + t[1] / 1e6;
}

function installSpeedballGI(opts) {
    opts = opts || {};
    const scene = opts.scene;
    if (!scene) throw new Error('installSpeedballGI: { scene } is required.');
    const onError = opts.onError == null ? null : opts.onError;
    if (onError !== null && typeof onError !== 'function') {
        throw new TypeError('installSpeedballGI: onError must be a function when provided.');
    }
    const camera = opts.camera || null;
    const gi = createProbeField({
        scene: scene,
        intensity: opts.intensity == null ? 10 : opts.intensity,
        hysteresis: opts.hysteresis == null ? null : opts.hysteresis,
        jitterMode: opts.jitterMode || 'gated',
        divisions: opts.divisions == null ? 16 : opts.divisions,
        rays: opts.rays == null ? 64 : opts.rays,
        cascades: opts.cascades == null ? 2 : opts.cascades,
        continuous: opts.continuous !== false,
        autoDetectChanges: opts.autoDetectChanges !== false,
        precision: opts.precision, traversal: opts.traversal, device: opts.device,
        roughReflections: opts.roughReflections, reflectionQuality: opts.reflectionQuality,
        reflectionSkyFallback: opts.reflectionSkyFallback, clusteredLighting: opts.clusteredLighting,
        addon: opts.addon, addonPath: opts.addonPath,
    });
    if (opts.enabled !== false) gi.setEnabled(true);

    let lastInteraction = now();
    let lastCam = null;
    const cameraMoved = (cam) => {
        if (!cam) return false;
        const e = cam.matrixWorld && cam.matrixWorld.elements ? Array.prototype.slice.call(cam.matrixWorld.elements) : null;
        if (!e) return false;
        const moved = !lastCam || e.some((v, i) => Math.abs(v - lastCam[i]) > 1e-7);
        lastCam = e;
        return moved;
    };
    let disposed = false;
    let updateFailureWarned = false;
    const mutate = (method) => function () { if (disposed) return undefined; return gi[method].apply(null, arguments); };
    const query = (method) => function () { return gi[method].apply(null, arguments); };
    const advanced = Object.freeze({
        node: gi.node,
        tick: mutate('tick'),
        resetFramePacing: mutate('resetFramePacing'),
        requestRebuild: mutate('requestRebuild'),
        setDebugKnobs: mutate('setDebugKnobs'),
        getDebugKnobs: query('getDebugKnobs'),
        debug: Object.freeze({
            state: query('_debugState'),
            read: query('_debugRead'),
            tickParams: query('_debugTickParams'),
        }),
    });

    const handle = {
        update: function (args) {
            if (disposed) return undefined;
            const a = args || {};
            const cam = a.camera === undefined ? camera : a.camera;
            const t = now();
            if (cameraMoved(cam)) lastInteraction = t;
            return gi.tick({ idleMs: t - lastInteraction, playing: a.playing === true }).then(function (result) {
                updateFailureWarned = false;
                return result;
            }).catch(function (error) {
                if (disposed || updateFailureWarned) return undefined;
                updateFailureWarned = true;
                if (onError) {
                    try { onError(error); } catch (cbError) { console.warn('SPEEDBALL GI onError callback failed:', cbError); }
                    return undefined;
                }
                console.warn('SPEEDBALL GI update failed:', error);
                return undefined;
            });
        },
        markInteraction: function () { if (!disposed) lastInteraction = now(); },
        advanced: advanced,
        dispose: function () {
            if (disposed) return;
            disposed = true;
            gi.dispose();
        },
        markMaterialsDirty: function () {},
    };
    const forwarded = ['setEnabled', 'setIntensity', 'setReflectionIntensity', 'setRoughnessLimit', 'setReflectionSkyFallback',
        'setChebyStrength', 'setNormalDetail', 'setClassifyStrength', 'setDivisions', 'setRays', 'setRayBudget',
        'setFilterStrength', 'setSmoothness', 'setHysteresis', 'setHysteresisNormalization', 'setJitterMode',
        'setNormalBias', 'setRadianceClamp', 'setDepthSharpness', 'setSky', 'setSkyIntensity', 'setChangeThreshold',
        'setSnapAmount', 'setFireflyClamp', 'forceLightingRefresh', 'setCascades', 'setContinuous',
        'setAutoDetectChanges', 'markTransformsDirty', 'markDeformsDirty', 'markMaterialValuesDirty', 'markTopologyDirty',
        'notifySceneChange', 'setBounds', 'setVolumes'];
    const queried = ['getEnabled', 'getIntensity', 'getReflectionIntensity', 'getRoughnessLimit', 'getReflectionSkyFallback',
        'hasRoughReflections', 'hasGlossyReflections', 'getReflectionQuality', 'getDivisions', 'getRays', 'getRayBudget',
        'getHysteresis', 'getHysteresisNormalization', 'getJitterMode', 'getSkyIntensity', 'getNirSensing', 'getNirGain',
        'getChangeThreshold', 'getSnapAmount', 'getFireflyClamp', 'getCascades', 'getContinuous', 'getAutoDetectChanges',
        'isSupported', 'hasData', 'getStats', 'getResolution', 'getBounds'];
    for (const m of forwarded) handle[m] = mutate(m);
    for (const m of queried) handle[m] = query(m);
    // NIR sensing flips the probes' NEE gate (the raster term is outside this build)
    handle.setNirSensing = function (on) { if (!disposed) gi.setNirSensing(on); };
    handle.setNirGain = function (g) { if (!disposed) gi.setNirGain(g); };
    return handle;
}

module.exports = { installSpeedballGI };
