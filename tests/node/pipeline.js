'use strict';
// Runs the reference's stereo -> filter -> fusion call sequence (reconstructor.ts:82-152 shape)
// through the Node host (lr3d.js over lr3d.node) on a job written by tests/test_node_host.py,
// and writes the packed depth maps, filtered maps and TSDF volume as raw little-endian files.
const fs = require('fs');
const path = require('path');
const lr3d = require(path.join(__dirname, '..', '..', 'luvas.io-3d_amd', 'node', 'lr3d.js'));

async function main() {
    const dir = process.argv[2];
    const job = JSON.parse(fs.readFileSync(path.join(dir, 'job.json'), 'utf8'));
    const ctx = new lr3d.Context(0);
    const input = {
        width: job.width,
        height: job.height,
        views: job.views,
        sources: job.sources,
        ranges: job.ranges,
        gray: async (layer) => new Uint8Array(fs.readFileSync(path.join(dir, `gray${layer}.u8`))),
        prior: () => null,
    };
    const run = await lr3d.STEREO_METHODS[job.method].create(ctx, input, { iterations: job.iterations, fineIterations: job.iterations });
    for (let layer = 0; layer < job.views.length; layer++) await run.solve(layer);
    const grid = new lr3d.BrickGrid(job.origin, job.voxel, job.extent);
    grid.mark(job.centre[0], job.centre[1], job.centre[2], job.radius);
    const layout = grid.build();
    // constructed as the reference does (tsdf-meshing.ts:71): the volume binds `stereo.filtered`
    const volume = new lr3d.HipTsdfVolume(ctx, grid, layout, run.stereo.filtered, job.width, job.height, job.truncation);
    for (let layer = 0; layer < job.views.length; layer++) {
        fs.writeFileSync(path.join(dir, `depth${layer}.u32`), Buffer.from((await run.stereo.depthMap(layer)).buffer));
        await run.stereo.filter(layer, job.sources[layer], job.needed);
        fs.writeFileSync(path.join(dir, `filtered${layer}.u32`), Buffer.from((await run.stereo.filteredMap(layer, job.sources[layer], job.needed)).buffer));
        await volume.integrate(Object.assign({}, job.views[layer], { maxDepth: job.ranges[layer][1], prior: null, confidence: 0 }));
    }
    // adjoints through the forward/backward pairs
    const opts = {
        sources: job.sources[0], minDepth: job.ranges[0][0], maxDepth: job.ranges[0][1], prior: null,
        geometric: false, iterations: job.iterations,
    };
    const ncc = new lr3d.NccCostFunction(run.stereo, 0, opts);
    const f = await ncc.forward(null);
    const g = await ncc.backward();
    fs.writeFileSync(path.join(dir, 'ncc_cost.f32'), Buffer.from(f.cost.buffer));
    fs.writeFileSync(path.join(dir, 'ncc_gd.f32'), Buffer.from(g.gradDepth.buffer));
    fs.writeFileSync(path.join(dir, 'ncc_gn.f32'), Buffer.from(g.gradNormal.buffer));
    fs.writeFileSync(path.join(dir, 'ncc_gp.f64'), Buffer.from(g.gradPose.buffer));
    const layers = job.views.map((v, i) => i);
    const ivs = job.views.map((v, i) => Object.assign({}, v, { maxDepth: job.ranges[i][1], prior: null, confidence: 0 }));
    const integration = new lr3d.TsdfIntegrationFunction(volume, layers, ivs, lr3d.MAP_FILTERED);
    const tb = await integration.backward(job.voxels, job.targets);
    fs.writeFileSync(path.join(dir, 'tsdf_bw.f32'), Buffer.from(tb.tsdf.buffer));
    fs.writeFileSync(path.join(dir, 'tsdf_gd.f32'), Buffer.from(tb.gradDepth.buffer));
    fs.writeFileSync(path.join(dir, 'tsdf_gp.f64'), Buffer.from(tb.gradPose.buffer));
    fs.writeFileSync(path.join(dir, 'losses.json'), JSON.stringify({ ncc: f.loss, tsdf: tb.loss }));
    // dense cloud of the filtered maps with the thumbnails written by the test
    const thumbs = layers.map((i) => ({ width: job.thumb[0], height: job.thumb[1],
        rgb: new Uint8Array(fs.readFileSync(path.join(dir, `thumb${i}.rgb`))) }));
    const cloud = await run.stereo.fuseDepthMaps(layers, thumbs, job.cell, job.cellBudget);
    for (const k of ['positions', 'normals', 'colors', 'support']) fs.writeFileSync(path.join(dir, `cloud_${k}.bin`), Buffer.from(cloud[k].buffer));
    fs.writeFileSync(path.join(dir, 'cloud.json'), JSON.stringify({ cell: cloud.cell }));
    const mesh = await volume.extractSurface(0.6, 0.01, 200);
    for (const k of ['positions', 'normals', 'indices']) fs.writeFileSync(path.join(dir, `mesh_${k}.bin`), Buffer.from(mesh[k].buffer));
    const { tsdf, weight } = await volume.read();
    fs.writeFileSync(path.join(dir, 'tsdf.f32'), Buffer.from(tsdf.buffer));
    fs.writeFileSync(path.join(dir, 'weight.f32'), Buffer.from(weight.buffer));
    fs.writeFileSync(path.join(dir, 'coords.i32'), Buffer.from(layout.coords.buffer));
    const solves = job.views.map((v, i) => [i, { sources: job.sources[i], minDepth: job.ranges[i][0],
        maxDepth: job.ranges[i][1], prior: null, geometric: false, iterations: job.iterations }]);
    // the rest of the C ABI through the Node host: allocation-free solves, launch shapes, timing,
    // counting, host-map integration and volume upload
    const extras = {};
    const before = [];
    for (let layer = 0; layer < job.views.length; layer++) before.push(await run.stereo.depthMap(layer));
    await run.stereo.solveBatchedNoalloc(solves);
    let same = true;
    for (let layer = 0; layer < job.views.length; layer++) {
        const after = await run.stereo.depthMap(layer);
        for (let i = 0; i < after.length; i++) if (after[i] !== before[layer][i]) { same = false; break; }
    }
    extras.noallocSame = same;
    ctx.setProfiling(true);
    await run.stereo.solve(0, solves[0][1]);
    extras.propagateLaunches = ctx.kernelStats('pm_propagate').launches;
    ctx.setProfiling(false);
    ctx.synchronize();
    extras.tiles = run.stereo.tiles();
    run.stereo.setTile('propagate', 16, 4);
    extras.tilesSet = run.stereo.tiles().propagate;
    run.stereo.setTile('propagate', extras.tiles.propagate[0], extras.tiles.propagate[1]);
    run.stereo.writeDepth(1, before[1]);
    const v2 = new lr3d.HipTsdfVolume(ctx, grid, layout, run.stereo.filtered, job.width, job.height, job.truncation);
    v2.setCounting(true);
    await v2.integrateLayers(layers, ivs, lr3d.MAP_FILTERED);
    extras.passCounts = v2.passCounts();
    extras.updateCount = v2.updateCount();
    extras.bricksPerBlock = await v2.autotune(layers, ivs, lr3d.MAP_FILTERED, 1);
    extras.xcdChunk = v2.xcdChunk;
    v2.setDepthCulling(false);
    v2.write(tsdf, weight);
    const back = await v2.read();
    extras.writeRoundTrip = back.tsdf.every((x, i) => Object.is(x, tsdf[i])) && back.weight.every((x, i) => Object.is(x, weight[i]));
    const v3 = new lr3d.HipTsdfVolume(ctx, grid, layout, run.stereo, job.width, job.height, job.truncation);
    const v4 = new lr3d.HipTsdfVolume(ctx, grid, layout, run.stereo, job.width, job.height, job.truncation);
    await v3.integrateHost(before[0], ivs[0]);
    await v4.integrateLayers([0], [ivs[0]], lr3d.MAP_RAW);
    const r3 = await v3.read(), r4 = await v4.read();
    extras.hostSame = r3.tsdf.every((x, i) => Object.is(x, r4.tsdf[i])) && r3.weight.every((x, i) => Object.is(x, r4.weight[i]));
    [v2, v3, v4].forEach((v) => v.destroy());
    fs.writeFileSync(path.join(dir, 'extras.json'), JSON.stringify(extras));
    // gradient refinement of planes and poses through the Node host (last: it moves the poses)
    const history = await lr3d.refine(run.stereo, solves, { steps: 5 });
    fs.writeFileSync(path.join(dir, 'refine.json'), JSON.stringify(history));
    // the coarse-to-fine driver over two coarse copies (half and quarter resolution)
    const grays = await Promise.all(job.views.map((v, i) => input.gray(i)));
    const levels = lr3d.pyramidLevels(run.stereo, grays, 2);
    const pyramid = await lr3d.refinePyramid(run.stereo, levels, solves, { coarseSteps: [2, 2], steps: 2 });
    fs.writeFileSync(path.join(dir, 'refine_pyramid.json'), JSON.stringify(pyramid));
    levels.forEach((l) => l.destroy());
    volume.destroy();
    run.destroy();
    ctx.destroy();
    console.log('node pipeline done', job.views.length, 'views', layout.count, 'bricks');
}

main().catch((e) => {
    console.error('node pipeline failed:', e.message);
    process.exit(1);
});
