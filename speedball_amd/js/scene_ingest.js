'use strict';
// Scene ingestion for the probe field: turns what a caller has into the reference's
// buffer layouts (js/spectral_scene.js:898-1366) through the native scene builder
// (sbgi_sb_* via the addon). three is not installed in Node, so three.js scenes are
// accepted by duck typing; plain typed-array descriptions and the reference's own
// prebuilt buffers are accepted as they are.
//
// Accepted inputs:
//   1. prebuilt reference buffers: { bvhNodes, triIndex, vertexData, triMaterial,
//      materials, instBase, instCount, tlasBase, tlasNodeCount, lights } — passed through;
//   2. plain description: { meshes: [{ positions, indices, normals?, uvs?, material? ,
//      materials?: per-triangle indices into `materials` }], materials?: [...],
//      instances: [{ mesh, matrix (16, column-major like Matrix4.elements), active? }],
//      lights: [{ type: 'point'|'directional'|'spot', position, target?, color, intensity,
//      distance?, decay?, angle?, penumbra?, emitterClass? }],
//      emitters?: [{ center, radius, area, radiance } or 17-float type-3 records] };
//   3. three-like scene graph: an object with traverse(cb) visiting objects that carry
//      isMesh / geometry.attributes.position / geometry.index / matrixWorld.elements /
//      material, and isPointLight / isDirectionalLight / isSpotLight lights; meshes with
//      userData.giEmitter also become emitter records (collectEmitterRecords).

const MAT_STRIDE = 28;
const LIGHT_STRIDE = 17;

function clamp(v, lo, hi) { return Math.min(hi, Math.max(lo, v)); }

function colorOf(c, dflt) {
    if (c == null) return dflt;
    if (Array.isArray(c) || ArrayBuffer.isView(c)) return [c[0], c[1], c[2]];
    if (typeof c === 'number') return [((c >> 16) & 255) / 255, ((c >> 8) & 255) / 255, (c & 255) / 255];
    if (typeof c.r === 'number') return [c.r, c.g, c.b];
    return dflt;
}

// materialToUber (js/spectral_scene.js:214-307) for map-free materials: 28 floats.
function uberRecord(m) {
    m = m || {};
    const r = new Float32Array(MAT_STRIDE);
    const color = colorOf(m.color, [1, 1, 1]);
    r[0] = color[0]; r[1] = color[1]; r[2] = color[2];
    r[3] = clamp(m.roughness == null ? 1 : m.roughness, 0, 1);
    r[4] = clamp(m.metalness == null ? 0 : m.metalness, 0, 1);
    const tr = clamp(m.transmission == null ? 0 : m.transmission, 0, 1);
    r[5] = tr;
    r[6] = Math.max(1, m.ior == null ? 1.5 : m.ior);
    const e = colorOf(m.emissive, [0, 0, 0]);
    const ei = m.emissiveIntensity == null ? 1 : m.emissiveIntensity;
    r[7] = e[0] * ei; r[8] = e[1] * ei; r[9] = e[2] * ei;
    const opacity = clamp(m.opacity == null ? 1 : m.opacity, 0, 1);
    r[10] = opacity;
    for (let i = 12; i < 17; i++) r[i] = -1;
    r[17] = 1; r[18] = 1; r[19] = 1;
    r[22] = clamp(Math.round(m.side == null ? 0 : m.side), 0, 2);
    const alphaTest = clamp(m.alphaTest == null ? 0 : m.alphaTest, 0, 1);
    r[23] = alphaTest;
    r[24] = -1; r[25] = -1;
    const constAlphaOk = opacity > 1e-4 && (alphaTest <= 0 || opacity >= alphaTest);
    r[26] = (constAlphaOk ? 2 : 0) | (tr < 0.5 ? 4 : 0);
    return r;
}

// collectLights records (js/spectral_scene.js:754-813): 17 floats.
function lightRecord(l) {
    const r = new Float32Array(LIGHT_STRIDE);
    const type = l.type || (l.isDirectionalLight ? 'directional' : l.isSpotLight ? 'spot' : 'point');
    r[0] = type === 'directional' ? 0 : type === 'spot' ? 2 : 1;
    const pos = worldPosition(l);
    r[1] = pos[0]; r[2] = pos[1]; r[3] = pos[2];
    let dir = [0, 0, -1];
    if (type !== 'point') {
        const tgt = l.target ? worldPosition(l.target) : [0, 0, 0];
        const d = [tgt[0] - pos[0], tgt[1] - pos[1], tgt[2] - pos[2]];
        const n = Math.hypot(d[0], d[1], d[2]) || 1;
        dir = [d[0] / n, d[1] / n, d[2] / n];
    }
    r[4] = dir[0]; r[5] = dir[1]; r[6] = dir[2];
    const c = colorOf(l.color, [1, 1, 1]);
    const I = l.intensity == null ? 1 : l.intensity;
    r[7] = c[0] * I; r[8] = c[1] * I; r[9] = c[2] * I;
    r[10] = type === 'directional' ? 0 : (l.distance || 0);
    r[11] = l.decay == null ? 2 : l.decay;
    if (type === 'spot') {
        const angle = l.angle == null ? Math.PI / 3 : l.angle;
        const pen = l.penumbra || 0;
        r[12] = Math.cos(angle);
        r[13] = Math.cos(angle * (1 - pen));
    } else {
        r[12] = -1; r[13] = -1;
    }
    r[16] = l.emitterClass === 'ir' || l.emitterClass === 4 ? 4 : 0;
    return r;
}

// collectEmitterRecords (js/spectral_scene.js:815-875): an emissive mesh as a type-3
// record {centre, radius (bounding sphere), surface radiance, range, A/4}. Accepts
// { center, radius, area, radiance } directly, or a three-like mesh whose geometry gives
// the area and the bounding sphere (box centre, farthest vertex) in world space.
function emitterRecord(e) {
    let center = e.center, radius = e.radius, area = e.area, rad = e.radiance;
    if (center == null && e.geometry && e.geometry.attributes && e.geometry.attributes.position) {
        const pos = e.geometry.attributes.position.array;
        const vc = pos.length / 3;
        const idx = e.geometry.index ? e.geometry.index.array : null;
        const w = worldMatrix(e);
        const tw = (i) => [w[0] * pos[3 * i] + w[4] * pos[3 * i + 1] + w[8] * pos[3 * i + 2] + w[12],
                           w[1] * pos[3 * i] + w[5] * pos[3 * i + 1] + w[9] * pos[3 * i + 2] + w[13],
                           w[2] * pos[3 * i] + w[6] * pos[3 * i + 1] + w[10] * pos[3 * i + 2] + w[14]];
        const n = idx ? idx.length : vc;
        area = 0;
        for (let i = 0; i + 2 < n; i += 3) {
            const a = tw(idx ? idx[i] : i), b = tw(idx ? idx[i + 1] : i + 1), c = tw(idx ? idx[i + 2] : i + 2);
            const u = [b[0] - a[0], b[1] - a[1], b[2] - a[2]], v = [c[0] - a[0], c[1] - a[1], c[2] - a[2]];
            area += 0.5 * Math.hypot(u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]);
        }
        const lo = [Infinity, Infinity, Infinity], hi = [-Infinity, -Infinity, -Infinity];
        for (let i = 0; i < vc; i++)
            for (let k = 0; k < 3; k++) { lo[k] = Math.min(lo[k], pos[3 * i + k]); hi[k] = Math.max(hi[k], pos[3 * i + k]); }
        const c = [(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2];
        let r2 = 0;
        for (let i = 0; i < vc; i++)
            r2 = Math.max(r2, (pos[3 * i] - c[0]) ** 2 + (pos[3 * i + 1] - c[1]) ** 2 + (pos[3 * i + 2] - c[2]) ** 2);
        const sc = Math.max(Math.hypot(w[0], w[1], w[2]), Math.hypot(w[4], w[5], w[6]), Math.hypot(w[8], w[9], w[10]));
        center = [w[0] * c[0] + w[4] * c[1] + w[8] * c[2] + w[12], w[1] * c[0] + w[5] * c[1] + w[9] * c[2] + w[13],
                  w[2] * c[0] + w[6] * c[1] + w[10] * c[2] + w[14]];
        radius = Math.sqrt(r2) * sc;
        if (rad == null && e.material) {
            const em = colorOf(e.material.emissive, [0, 0, 0]);
            const ei = e.material.emissiveIntensity == null ? 1 : e.material.emissiveIntensity;
            const scale = e.userData && Number.isFinite(e.userData.giEmitterScale) ? e.userData.giEmitterScale : 1;
            rad = [em[0] * ei * scale, em[1] * ei * scale, em[2] * ei * scale];
        }
    }
    rad = colorOf(rad, [1, 1, 1]);
    const r = new Float32Array(LIGHT_STRIDE);
    const proj = area / 4;
    r[0] = 3;
    r[1] = center[0]; r[2] = center[1]; r[3] = center[2];
    r[4] = 0; r[5] = 0; r[6] = -1;
    r[7] = rad[0]; r[8] = rad[1]; r[9] = rad[2];
    r[10] = Math.sqrt(Math.max(0, rad[0], rad[1], rad[2]) * proj / 0.005);
    r[11] = 2; r[12] = -1; r[13] = -1;
    r[14] = radius; r[15] = proj; r[16] = 0;
    return r;
}

function worldPosition(o) {
    if (Array.isArray(o.position) || ArrayBuffer.isView(o.position)) return [o.position[0], o.position[1], o.position[2]];
    if (o.matrixWorld && o.matrixWorld.elements) {
        const e = o.matrixWorld.elements;
        return [e[12], e[13], e[14]];
    }
    if (o.position && typeof o.position.x === 'number') return [o.position.x, o.position.y, o.position.z];
    return [0, 0, 0];
}

// world matrix as a column-major Float64Array(16) (three.js Matrix4.elements order)
function worldMatrix(o) {
    const src = o.matrix || (o.matrixWorld && o.matrixWorld.elements) || null;
    const out = new Float64Array(16);
    if (!src) { out[0] = out[5] = out[10] = out[15] = 1; return out; }
    for (let i = 0; i < 16; i++) out[i] = src[i];
    return out;
}

function isPrebuilt(s) {
    return s && s.bvhNodes && s.triIndex && s.vertexData && s.materials && s.instCount != null;
}

function describeThreeScene(scene) {
    const desc = { meshes: [], materials: [], instances: [], lights: [], objects: [], emitters: [] };
    const meshOfGeometry = new Map();
    const matIndex = new Map();
    scene.traverse(function (o) {
        if (o.visible === false) return;
        if (o.isMesh && o.geometry && o.geometry.attributes && o.geometry.attributes.position) {
            let mesh = meshOfGeometry.get(o.geometry);
            const mat = Array.isArray(o.material) ? o.material[0] : o.material;
            let mi = matIndex.get(mat);
            if (mi == null) {
                mi = desc.materials.length;
                desc.materials.push(mat || {});
                matIndex.set(mat, mi);
            }
            if (mesh == null) {
                const g = o.geometry;
                const pos = g.attributes.position.array;
                const vcount = pos.length / 3;
                let idx = g.index ? g.index.array : null;
                if (!idx) { idx = new Uint32Array(vcount); for (let i = 0; i < vcount; i++) idx[i] = i; }
                mesh = desc.meshes.length;
                desc.meshes.push({
                    positions: pos, indices: idx, materialIndex: mi, geometry: g,
                    version: g.attributes.position.version || 0,
                });
                meshOfGeometry.set(g, mesh);
            }
            desc.instances.push({ mesh: mesh, matrix: worldMatrix(o), active: true, object: o });
            desc.objects.push(o);
            if (o.userData && o.userData.giEmitter === true) desc.emitters.push(o);
        } else if (o.isPointLight || o.isDirectionalLight || o.isSpotLight) {
            desc.lights.push(o);
        }
    });
    return desc;
}

// Build the reference buffers. Returns { arrays, meshes, instances, builder } where
// arrays are the reference layouts (the sbgi_scene_view fields) and instances[i] keeps
// the builder instance id, its mesh and its source object for later transform updates.
function buildScene(addon, input) {
    if (isPrebuilt(input)) {
        return { arrays: input, meshes: [], instances: [], builder: null, prebuilt: true };
    }
    const desc = (typeof input.traverse === 'function') ? describeThreeScene(input) : input;
    const b = addon.sbCreate();
    const matIds = [];
    const mats = desc.materials || [];
    for (let i = 0; i < mats.length; i++) matIds.push(addon.sbAddMaterial(b, uberRecord(mats[i])));
    const meshes = [];
    let vertBase = 0;
    for (let m = 0; m < desc.meshes.length; m++) {
        const mesh = desc.meshes[m];
        const pos = mesh.positions instanceof Float32Array ? mesh.positions : new Float32Array(mesh.positions);
        const idx = mesh.indices instanceof Uint32Array ? mesh.indices : new Uint32Array(mesh.indices);
        const tcount = idx.length / 3;
        let triMat;
        if (mesh.materials) {
            triMat = new Uint32Array(tcount);
            for (let t = 0; t < tcount; t++) triMat[t] = matIds[mesh.materials[t]];
        } else {
            let id;
            if (mesh.materialIndex != null) id = matIds[mesh.materialIndex];
            else { id = addon.sbAddMaterial(b, uberRecord(mesh.material)); matIds.push(id); }
            triMat = new Uint32Array(tcount).fill(id);
        }
        const nrm = mesh.normals ? Float32Array.from(mesh.normals) : null;
        const uvs = mesh.uvs ? Float32Array.from(mesh.uvs) : null;
        addon.sbAddMesh(b, pos, nrm, uvs, idx, triMat);
        meshes.push({ vertBase: vertBase, vcount: pos.length / 3, source: mesh });
        vertBase += pos.length / 3;
    }
    const instances = [];
    for (let i = 0; i < desc.instances.length; i++) {
        const inst = desc.instances[i];
        const id = addon.sbAddInstance(b, inst.mesh, worldMatrix(inst), inst.active !== false);
        instances.push({ id: id, mesh: inst.mesh, object: inst.object || inst, slot: -1 });
    }
    const lights = desc.lights || [];
    for (let i = 0; i < lights.length; i++) addon.sbAddLight(b, lightRecord(lights[i]));
    // emitters follow the analytic lights in the arena (the builder orders them)
    const emitters = desc.emitters || [];
    for (let i = 0; i < emitters.length; i++) {
        const e = emitters[i];
        addon.sbAddLight(b, e instanceof Float32Array && e.length === LIGHT_STRIDE ? e : emitterRecord(e));
    }
    const arrays = addon.sbFinalize(b);
    // builder instance id -> TLAS slot
    const order = arrays.tlasOrder;
    const slotOf = new Map();
    for (let s = 0; s < order.length; s++) slotOf.set(order[s], s);
    for (const inst of instances) inst.slot = slotOf.get(inst.id);
    return { arrays: arrays, meshes: meshes, instances: instances, builder: b, prebuilt: false, desc: desc };
}

// Bounds of the traced soup: every instance's world-space vertex AABB.
function sceneBounds(built) {
    const a = built.arrays;
    const lo = [Infinity, Infinity, Infinity], hi = [-Infinity, -Infinity, -Infinity];
    const tl = a.materials;
    // TLAS root node = union of all instance world AABBs (12-float node at tlasBase)
    if (a.tlasNodeCount > 0) {
        for (let k = 0; k < 3; k++) { lo[k] = tl[a.tlasBase + k]; hi[k] = tl[a.tlasBase + 3 + k]; }
    }
    return { min: lo, max: hi };
}

module.exports = {
    buildScene, uberRecord, lightRecord, emitterRecord, worldMatrix, sceneBounds, MAT_STRIDE, LIGHT_STRIDE,
};
