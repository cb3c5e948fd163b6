'use strict';
// Drives the JS facade (webgpu-bcmc_amd/js/index.js) the way the reference's
// app drives its classes (js/bcmc.js:45-47,352-375) and writes the results for
// tests/test_gpu_node.py to compare with the oracle.
// Usage: node facade_parity.js <dir>   (reads <dir>/params.json, <dir>/stream.bin)
const fs = require('fs');
const path = require('path');
const bcmc = require(path.join(__dirname, '..', '..', 'webgpu-bcmc_amd', 'js', 'index.js'));

async function main() {
    const dir = process.argv[2];
    const p = JSON.parse(fs.readFileSync(path.join(dir, 'params.json')));
    const stream = new Uint8Array(fs.readFileSync(path.join(dir, 'stream.bin')));
    const device = await bcmc.requestDevice(0);
    const out = {counts: [], perfKeys: null};

    // ExclusiveScanPipeline / ExclusiveScanner
    const scanIn = Uint32Array.from(p.scan_input);
    const pipe = new bcmc.ExclusiveScanPipeline(device);
    const scanner = pipe.prepareInput(scanIn);
    out.scan_total = await scanner.scan(scanIn.length);
    out.scan = Array.from(scanner.inputBuf.read(Uint32Array, scanIn.length));

    // StreamCompactor
    const flags = Uint32Array.from(p.flags);
    const offs = pipe.prepareInput(flags);
    const nflag = await offs.scan(flags.length);
    const ids = device.createBuffer({size: 4 * Math.max(nflag, 1)});
    await new bcmc.StreamCompactor(device).compactActiveIDs(
        flags.length, device.bufferFrom(flags), offs.inputBuf, ids);
    out.compact = Array.from(ids.read(Uint32Array, nflag));

    // RadixSorter (reverse)
    const sorter = new bcmc.RadixSorter(device);
    const n = p.sort_keys.length, a = sorter.getAlignedSize(n);
    const k = new Uint32Array(a), v = new Uint32Array(a);
    k.set(p.sort_keys);
    for (let i = 0; i < n; ++i) v[i] = i;
    const kb = device.bufferFrom(k), vb = device.bufferFrom(v);
    await sorter.sort(kb, vb, n, true);
    out.sorted = Array.from(vb.read(Uint32Array, n));

    // CompressedMarchingCubes sweep
    const mc = new bcmc.CompressedMarchingCubes(device);
    await mc.setCompressedVolume(stream, p.rate, p.dims, [1, 1, 1], p.cache);
    const perf = {};
    for (let i = 0; i < p.isos.length; ++i) {
        const nv = await mc.computeSurface(p.isos[i], perf);
        const verts = mc.vertexBuffer.read(Float32Array, 4 * nv);
        fs.writeFileSync(path.join(dir, `verts_${i}.bin`), Buffer.from(verts.buffer));
        const st = mc.lruCache.state();
        fs.writeFileSync(path.join(dir, `slots_${i}.bin`), Buffer.from(st.itemSlot.buffer));
        out.counts.push({
            numVertices: nv,
            numActiveBlocks: mc.numActiveBlocks,
            numBlocksWithVertices: mc.numBlocksWithVertices,
            newDecompressed: mc.newDecompressed,
            cacheSize: mc.lruCache.cacheSize
        });
    }
    out.perfKeys = Object.keys(perf).sort();
    out.perfLen = perf.numVertices.length;
    out.memory = mc.reportMemoryUse()[3];

    // persisted block ranges: without and with the tile ranges, loaded into
    // new objects in place of the load-time decode; the second object's
    // untimed queries replay a captured graph of the query chain
    const persisted = {block: mc.blockRanges(false), tile: mc.blockRanges(true)};
    fs.writeFileSync(path.join(dir, 'ranges.bin'), Buffer.from(persisted.block.buffer));
    out.persisted = {};
    for (const name of Object.keys(persisted)) {
        device.setGraphReplay(name === 'tile');
        const m2 = new bcmc.CompressedMarchingCubes(device);
        await m2.setCompressedVolume(stream, p.rate, p.dims, [1, 1, 1], p.cache, persisted[name]);
        out.persisted[name] = [];
        for (let i = 0; i < p.isos.length; ++i) {
            const nv = await m2.computeSurface(p.isos[i]);
            const verts = m2.vertexBuffer.read(Float32Array, 4 * nv);
            fs.writeFileSync(path.join(dir, `pverts_${name}_${i}.bin`), Buffer.from(verts.buffer));
            out.persisted[name].push({numVertices: nv, newDecompressed: m2.newDecompressed});
        }
        m2.destroy();
    }
    device.setGraphReplay(false);
    await mc.lruCache.reset();
    out.after_reset_age0 = mc.lruCache.state().age[0];
    mc.destroy();
    device.destroy();
    fs.writeFileSync(path.join(dir, 'out.json'), JSON.stringify(out));
    console.log('facade_parity: ok');
}

main().catch((e) => {
    console.error('facade_parity failed:', e);
    process.exit(1);
});
