// The engine from Node: three volumes (<dir>/vol_<i>.bin, params.json), one
// batched step, then a served trace; writes counts and vertices for the test.
'use strict';
const fs = require('fs');
const path = require('path');
const bcmc = require(path.join(__dirname, '..', '..', 'webgpu-bcmc_amd', 'js', 'index.js'));

(async () => {
    const dir = process.argv[2];
    const p = JSON.parse(fs.readFileSync(path.join(dir, 'params.json')));
    const dev = await bcmc.requestDevice();
    const eng = new bcmc.Engine(dev, p.poolBytes);
    const ids = p.vols.map((v, i) =>
        eng.addVolume(new Uint8Array(fs.readFileSync(path.join(dir, `vol_${i}.bin`))), v.rate, v.dims));
    const res = await eng.run(p.qv.map((k) => ids[k]), p.qi);
    const total = res.reduce((a, r) => a + r.numVertices, 0);
    const verts = eng.vertexBuffer.read(Float32Array, 4 * total);
    fs.writeFileSync(path.join(dir, 'verts.bin'), Buffer.from(verts.buffer));
    // world-space positions of the first query's vertices (before serve()
    // replaces the vertex buffer with its last step's)
    const q0 = res[0];
    const w = await bcmc.worldPositions(dev, eng.vertexBuffer, q0.numVertices, p.vols[p.qv[0]].dims,
                                        p.scale, q0.vertexOffset);
    fs.writeFileSync(path.join(dir, 'world.bin'), Buffer.from(w.read(Float32Array, 3 * q0.numVertices).buffer));
    const served = await eng.serve(p.arrivals, p.qv.map((k) => ids[k]), p.qi, {maxBatch: 4});
    const stats = eng.stats();
    eng.destroy();
    // GPU encoder on a float field
    const raw = fs.readFileSync(path.join(dir, 'field.bin'));
    const field = new Float32Array(raw.byteLength / 4);
    new Uint8Array(field.buffer).set(raw);
    const enc = await bcmc.zfpEncode(dev, dev.bufferFrom(field), p.fieldDims, p.fieldRate);
    fs.writeFileSync(path.join(dir, 'encoded.bin'), Buffer.from(enc.read(Uint8Array, enc.size).buffer));
    fs.writeFileSync(path.join(dir, 'out.json'), JSON.stringify({res, served, stats}));
})().catch((e) => {
    console.error(e);
    process.exit(1);
});
