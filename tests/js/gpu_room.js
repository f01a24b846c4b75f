'use strict';
// GPU run of the JS shim through the real addon: loads a prebuilt scene (the
// reference buffers written by the Python test), places one manual volume, runs two
// updates and writes the irradiance state and atlas for comparison.
//   node gpu_room.js <dir>
const fs = require('fs');
const path = require('path');
const shim = require(path.join(__dirname, '..', '..', 'speedball_amd', 'js'));

const dir = process.argv[2];
const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
function load(name, T) {
    const b = fs.readFileSync(path.join(dir, name + '.bin'));
    return new T(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
}
const scene = {
    bvhNodes: load('bvhNodes', Uint32Array), triIndex: load('triIndex', Uint32Array),
    vertexData: load('vertexData', Float32Array), triMaterial: load('triMaterial', Uint32Array),
    materials: load('materials', Float32Array), lights: load('lights', Float32Array),
    instBase: meta.instBase, instCount: meta.instCount, tlasBase: meta.tlasBase, tlasNodeCount: meta.tlasNodeCount,
};

(async () => {
    console.log('step: create field');
    const gi = shim.createProbeField({ scene: scene, rays: meta.rays, precision: meta.precision, traversal: meta.traversal });
    gi.setVolumes({ box: { min: meta.gridMin, max: meta.gridMax }, res: meta.res });
    gi.setHysteresis(meta.hysteresis);
    gi.setEnabled(true);
    for (let i = 0; i < 2; i++) {
        const r = await gi.tick({});
        console.log('step: tick', i, 'updated', r.updated, 'ms', r.stats.msTotal.toFixed(3));
    }
    const irr = await gi._debugRead('irr');
    fs.writeFileSync(path.join(dir, 'js_irr.bin'), Buffer.from(irr.buffer));
    const stats = gi.getStats();
    fs.writeFileSync(path.join(dir, 'js_stats.json'), JSON.stringify(stats));
    gi.dispose();
    console.log('done');
})().catch((e) => { console.error(e && e.stack || e); process.exit(1); });
