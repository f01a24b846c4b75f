'use strict';
// Loads the N-API addon (speedball_amd/napi/sbgi.node) that maps the C ABI of
// include/sbgi.h. There is no fallback: without the addon the shim throws, because a
// probe field without the HIP library has nothing to run on.
const path = require('path');

let cached = null;

function loadAddon(explicitPath) {
    if (explicitPath) return require(explicitPath);
    if (cached) return cached;
    const p = path.join(__dirname, '..', 'napi', 'sbgi.node');
    try {
        cached = require(p);
    } catch (e) {
        const err = new Error('speedball_amd: the native addon ' + p + ' is not built or cannot load ' +
            '(build it with __graft_entry__.build()): ' + e.message);
        err.code = 'SBGI_ADDON_MISSING';
        throw err;
    }
    return cached;
}

// Status codes and buffer ids of include/sbgi.h.
const STATUS = Object.freeze({
    OK: 0, E_INVALID: -1, E_LAYOUT: -2, E_CAPACITY: -3, E_DEVICE: -4, E_STATE: -5, E_TIMEOUT: -6, E_NOMEM: -7,
});
const BUF = Object.freeze({
    BVH_NODES: 0, TRI_INDEX: 1, VERTEX_DATA: 2, TRI_MATERIAL: 3, MATERIALS: 4, LIGHTS: 5,
    IRR_STATE: 10, DEPTH_STATE: 11, META: 12, IRR_ATLAS: 13, DEPTH_ATLAS: 14, RAY_DATA: 15, HIT_IDS: 16,
    NODES_REFIT: 17, TRI_PACKED: 18, INST_BOUNDS: 19, TLAS_SAH: 20, EMITTER_VIS: 21, LIGHT_GRID: 22,
    ROUGH_STATE: 23, ROUGH_ATLAS: 24, GLOSSY_STATE: 25, GLOSSY_WEIGHT: 26, GLOSSY_ATLAS: 27, TLAS_OCT: 28,
    SHADOW_VIS: 29,
});
// reflectionQuality tiers (js/gi_probes.js:60-65) -> SBGI_REFLECT_*
const REFLECTION = Object.freeze({ off: 0, rough: 1, high: 2, ultra: 3 });
const BLEND = Object.freeze({ fp32: 0, bf16: 1, bf16x3: 2, fp16: 3 });
const TRAVERSAL = Object.freeze({ compat: 0, ordered: 1, watertight: 2 });

module.exports = { loadAddon, STATUS, BUF, BLEND, TRAVERSAL, REFLECTION };
