// Drives the isovalue benchmarks.  argv[2]: output directory.  argv[3]:
// 'mock' runs the benchmark loop on a stand-in pipeline (no GPU); 'gpu' runs a
// sweepUp on a real CompressedMarchingCubes over <dir>/stream.bin.
'use strict';
const fs = require('fs');
const path = require('path');
const js = path.join(__dirname, '..', '..', 'webgpu-bcmc_amd', 'js');

async function mock(dir) {
    const b = require(path.join(js, 'benchmark.js'));
    const calls = [];
    const mc = {
        lruCache: {reset: async () => calls.push('reset')},
        computeSurface: async (iso, perf) => {
            calls.push(iso);
            if (perf.numVertices === undefined) perf.numVertices = [];
            perf.numVertices.push(7);
            return 7;
        },
        reportMemoryUse: () => ['a', 'b', 'c', {mc: {total: 1}}],
    };
    const out = {};
    for (const kind of ['sweepUp', 'sweepDown']) {
        calls.length = 0;
        const r = await b.runIsovalueBenchmark(mc, kind, [0.1, 0.9], {datasetName: 'skull'});
        out[kind] = {calls: calls.slice(), fileName: r.fileName, perf: r.perfResults};
    }
    let seed = 1;
    const lcg = () => { seed = (seed * 48271) % 2147483647; return seed / 2147483647; };
    calls.length = 0;
    const r = await b.runIsovalueBenchmark(mc, 'random', [0.25, 0.75], {random: lcg, enableCache: false});
    out.random = {calls: calls.slice(), fileName: r.fileName, perf: r.perfResults};
    let threw = false;
    try { await b.runIsovalueBenchmark(mc, 'zigzag', [0, 1]); } catch (e) { threw = true; }
    out.unknownThrows = threw;
    out.alias = b.SweepIsovalueBenchark === b.SweepIsovalueBenchmark;
    fs.writeFileSync(path.join(dir, 'out.json'), JSON.stringify(out));
}

async function gpu(dir) {
    const bcmc = require(path.join(js, 'index.js'));
    const p = JSON.parse(fs.readFileSync(path.join(dir, 'params.json')));
    const dev = await bcmc.requestDevice();
    const mc = new bcmc.CompressedMarchingCubes(dev);
    const stream = new Uint8Array(fs.readFileSync(path.join(dir, 'stream.bin')));
    await mc.setCompressedVolume(stream, p.rate, p.dims, [1, 1, 1], p.cache);
    await mc.computeSurface(0.5);  // the cache holds blocks before the benchmark's reset
    const r = await bcmc.runIsovalueBenchmark(mc, 'sweepUp', p.range, {datasetName: 'blob'});
    fs.writeFileSync(path.join(dir, 'out.json'), JSON.stringify({fileName: r.fileName, perf: r.perfResults}));
    mc.destroy();
}

(process.argv[3] === 'gpu' ? gpu : mock)(process.argv[2]).catch((e) => {
    console.error(e);
    process.exit(1);
});
