'use strict';
// Runs a batch through the Node facade on the GPU and prints the results
// as JSON: `node gpu_batch.js case.json` (one half-sweep per relaxation
// launch for the implicit steps, persistent launches for the solve). The
// case holds the grids, their Mach numbers and boundary kinds (an H grid's
// inflow iMin or a C grid's axis); the calling test compares the states
// and the solve with the Python facade on the same batch.
const fs = require('fs');
const bs = require('..');

(async () => {
  const c = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  const device = new bs.HipDevice(0);
  const frees = c.cases.map((k) => bs.freeStreamAt(k.mach, 1.4));
  const configs = c.cases.map((k, n) => {
    const inflow = { kind: 'inflow', state: frees[n].state };
    return {
      grid: { ni: k.ni, nj: k.nj, x: Float64Array.from(k.x), y: Float64Array.from(k.y) },
      geometry: 'axisymmetric',
      gamma: 1.4,
      flux: 'hllc',
      limiter: 'vanAlbada',
      reconstruction: 'sharedScaled',
      timeIntegrator: 'ssprk3',
      cfl: 0.5,
      timeStepping: 'local',
      boundaries: { iMin: k.axis ? { kind: 'wall' } : inflow, iMax: { kind: 'outflow' }, jMin: { kind: 'wall' }, jMax: inflow },
    };
  });
  const out = {};
  // Implicit steps of every case, the states read back.
  const batch = await bs.GpuBatch2D.create(device, configs);
  batch.setUniform(frees.map((f) => f.state));
  // One half-sweep per relaxation launch: the states do not change.
  batch.setRelaxFusion(1);
  await batch.setImplicitCfl(1);
  await batch.iterateImplicit(c.steps, { cfl: 10, sweeps: 2 });
  out.rho = [];
  for (let k = 0; k < batch.count; k++) out.rho.push(Array.from((await batch.readState(k)).rho));
  out.norms = Array.from(await batch.densityResidualNorms());
  out.drags = Array.from(await batch.forebodyDrags(frees, 0.5));
  // Overlapping calls are refused, as on a single solver.
  const running = batch.iterate(1);
  try {
    await batch.iterate(1);
    out.overlap = 'allowed';
  } catch (e) {
    out.overlap = e.message;
  }
  await running;
  // Coarse fields for every case, interpolated on the GPU.
  if (c.coarse) {
    batch.setCoarse(c.coarse, c.scales);
    out.coarseRho = [];
    for (let k = 0; k < batch.count; k++) out.coarseRho.push(Array.from((await batch.readState(k)).rho));
  }
  batch.destroy();
  // The stopping rule, explicit steps from the free stream.
  const solveBatch = await bs.GpuBatch2D.create(device, configs);
  solveBatch.setUniform(frees.map((f) => f.state));
  // Persistent launches: the same states, so the same outcomes.
  solveBatch.setPersistent(2);
  out.solve = await bs.solveBatch(solveBatch, frees, 0.5, { chunk: 240, maxSteps: c.maxSteps });
  solveBatch.destroy();
  // faceStates sides are refused.
  try {
    await bs.GpuBatch2D.create(device, [Object.assign({}, configs[0], {
      boundaries: Object.assign({}, configs[0].boundaries, { jMax: { kind: 'faceStates', states: [] } }),
    })]);
    out.faceStates = 'allowed';
  } catch (e) {
    out.faceStates = e.message;
  }
  device.destroy();
  console.log(JSON.stringify(out));
})().catch((e) => {
  console.error(e);
  process.exit(1);
});
