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
    const served = await eng.serve(p.arrivals, p.qv.map((k) => ids[k]), p.qi, {maxBatch: 4});
    const stats = eng.stats();
    eng.destroy();
    fs.writeFileSync(path.join(dir, 'out.json'), JSON.stringify({res, served, stats}));
})().catch((e) => {
    console.error(e);
    process.exit(1);
});
