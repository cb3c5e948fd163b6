// The async facade keeps the event loop running: a setInterval timer ticks
// while Engine.serve (requests arriving over about a second), a
// computeSurface and an LRUCache.update are pending.  Also fills an LRU
// perfTracker.  Writes out.json for the test.
'use strict';
const fs = require('fs');
const path = require('path');
const bcmc = require(path.join(__dirname, '..', '..', 'webgpu-bcmc_amd', 'js', 'index.js'));

async function ticksWhile(promise) {
    let ticks = 0;
    const timer = setInterval(() => { ticks += 1; }, 5);
    const t0 = Date.now();
    const value = await promise;
    const ms = Date.now() - t0;
    clearInterval(timer);
    return {value, ticks, ms};
}

(async () => {
    const dir = process.argv[2];
    const p = JSON.parse(fs.readFileSync(path.join(dir, 'params.json')));
    const dev = await bcmc.requestDevice();
    const vol = new Uint8Array(fs.readFileSync(path.join(dir, 'vol.bin')));
    // served trace: arrivals spread over `span` seconds
    const eng = new bcmc.Engine(dev, p.poolBytes);
    const id = eng.addVolume(vol, p.rate, p.dims);
    const served = await ticksWhile(
        eng.serve(p.arrivals, p.arrivals.map(() => id), p.isos, {maxBatch: 8}));
    eng.destroy();
    // computeSurface
    const mc = new bcmc.CompressedMarchingCubes(dev);
    await mc.setCompressedVolume(vol, p.rate, p.dims, [1, 1, 1], p.cache);
    const surf = await ticksWhile(mc.computeSurface(p.isos[0]));
    // LRUCache.update with a perfTracker, twice (new items, then none)
    const lru = new bcmc.LRUCache(dev, null, null, p.lruInitial, 256, p.lruTotal);
    const needed = new Uint32Array(p.lruTotal);
    for (let i = 0; i < p.lruTotal; i += 3) needed[i] = 1;
    const perf = {};
    const upd = await ticksWhile(lru.update(dev.bufferFrom(needed), perf));
    await lru.update(dev.bufferFrom(needed), perf);
    lru.destroy();
    // operations submitted without awaiting run in call order
    const order = [];
    const a = mc.computeSurface(p.isos[1]).then((n) => order.push(['a', n]));
    const b = mc.computeSurface(p.isos[2]).then((n) => order.push(['b', n]));
    await Promise.all([a, b]);
    fs.writeFileSync(path.join(dir, 'out.json'), JSON.stringify({
        serveTicks: served.ticks, serveMs: served.ms, served: served.value,
        surfaceVertices: surf.value, updateNew: upd.value[0], perf, order,
    }));
})().catch((e) => {
    console.error(e);
    process.exit(1);
});
