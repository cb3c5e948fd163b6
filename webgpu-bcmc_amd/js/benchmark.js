'use strict';
// Isovalue benchmarks of the reference app (js/run_benchmark.js:1-47) and the
// loop that drives them (js/bcmc.js:319-370), without the page.  The
// isovalue "slider" is any object with a numeric `value` (the page's range
// input snaps values to its step; this one does not).  Each benchmark sets
// 100 isovalues; the loop resets the LRU cache first, recomputes the surface
// whenever the value changed, and collects the perfTracker the page saves as
// `bcmc-explicit-<dataset>-<benchmark>.json`.

const kSteps = 100;

// isovalue = U[range) from `random` (the page uses Math.random)
class RandomIsovalueBenchmark {
    constructor(isovalueSlider, perfResults, range, random = Math.random) {
        this.iteration = 0;
        this.isovalueSlider = isovalueSlider;
        this.perfResults = perfResults;
        this.range = range;
        this.random = random;
        this.name = 'random';
    }
    run() {
        if (this.iteration === kSteps) return false;
        const span = this.range[1] - this.range[0];
        this.isovalueSlider.value = this.random() * span + this.range[0];
        this.iteration += 1;
        return true;
    }
}

// isovalues range[0] + k (range[1] - range[0]) / 100, k = 1..100, accumulated
// step by step (sweepUp), or down from range[1] (sweepDown)
class SweepIsovalueBenchmark {
    constructor(isovalueSlider, perfResults, range, sweepUp) {
        this.iteration = 0;
        this.isovalueSlider = isovalueSlider;
        this.perfResults = perfResults;
        this.range = range;
        this.sweepUp = sweepUp;
        this.currentValue = sweepUp ? range[0] : range[1];
        this.name = sweepUp ? 'sweepUp' : 'sweepDown';
    }
    run() {
        if (this.iteration === kSteps) return false;
        const step = (this.range[1] - this.range[0]) / kSteps;
        this.currentValue += this.sweepUp ? step : -step;
        this.isovalueSlider.value = this.currentValue;
        this.iteration += 1;
        return true;
    }
}

// Runs one benchmark ('random', 'sweepUp' or 'sweepDown') on a
// CompressedMarchingCubes with its volume set.  opts: datasetName (file name),
// random (source for 'random'), enableCache (false: reset the LRU before
// every step, as the page's cache checkbox does).  Returns the perf results
// ({isovalue, totalTime (ms), memUse, and the per-stage keys computeSurface
// adds}) and the file name the page would save them under.
async function runIsovalueBenchmark(compressedMC, kind, range, opts = {}) {
    const perfResults = {isovalue: [], totalTime: [], memUse: []};
    await compressedMC.lruCache.reset();
    const slider = {value: NaN};
    let bench;
    if (kind === 'random')
        bench = new RandomIsovalueBenchmark(slider, perfResults, range, opts.random || Math.random);
    else if (kind === 'sweepUp' || kind === 'sweepDown')
        bench = new SweepIsovalueBenchmark(slider, perfResults, range, kind === 'sweepUp');
    else
        throw new Error(`unknown benchmark '${kind}' (random, sweepUp or sweepDown)`);
    const enableCache = opts.enableCache !== false;
    let current = NaN;
    while (bench.run()) {
        if (!enableCache) await compressedMC.lruCache.reset();
        if (slider.value !== current) {
            current = slider.value;
            const start = process.hrtime.bigint();
            await compressedMC.computeSurface(current, perfResults);
            const end = process.hrtime.bigint();
            perfResults.isovalue.push(current);
            perfResults.totalTime.push(Number(end - start) / 1e6);
            perfResults.memUse.push(compressedMC.reportMemoryUse()[3]);
        }
    }
    const dataset = opts.datasetName || 'volume';
    return {perfResults, fileName: `bcmc-explicit-${dataset}-${bench.name}.json`};
}

module.exports = {
    RandomIsovalueBenchmark,
    SweepIsovalueBenchmark,
    // the reference's spelling (js/run_benchmark.js:19)
    SweepIsovalueBenchark: SweepIsovalueBenchmark,
    runIsovalueBenchmark,
};
