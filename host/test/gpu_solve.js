'use strict';
// Runs one case through the Node facade on the GPU and prints the results
// as JSON: `node gpu_solve.js case.json`. The case file holds the grid,
// the configuration, the free stream and the step counts; the calling test
// compares the output with the FP64 oracle on the same case.
const fs = require('fs');
const bs = require('..');

(async () => {
  const c = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  const free = bs.freeStreamAt(c.mach, c.config.gamma);
  const config = Object.assign({}, c.config, {
    grid: { ni: c.ni, nj: c.nj, x: Float64Array.from(c.x), y: Float64Array.from(c.y) },
  });
  for (const side of ['iMin', 'jMax']) config.boundaries[side].state = free.state;
  const device = new bs.HipDevice(0);
  const out = {};
  // The steady-engine adapter, as verifiedForebodyDrag drives it.
  const engine = await bs.hipEngine(device)(config, free.state);
  await engine.iterate(c.steps);
  const wall = await engine.wallFluxes();
  out.drag = bs.forebodyPressureDrag(wall, free, 0.5);
  const cells = await engine.cells();
  out.cell0 = cells(0, 0);
  engine.destroy();
  // The engine itself: overlapping calls are refused, the drag reduced on
  // the GPU equals the one from the wall fluxes.
  const gpu = await bs.GpuEuler2D.create(device, config);
  gpu.setInitialCondition(() => free.state);
  const running = gpu.iterate(c.steps);
  try {
    await gpu.iterate(1);
    out.overlap = 'allowed';
  } catch (e) {
    out.overlap = e.message;
  }
  await running;
  out.gpuDrag = await gpu.forebodyDrag(free, 0.5);
  out.residual = await gpu.densityResidualNorm();
  const state = await gpu.readState();
  out.steps = state.steps;
  out.rhoSum = state.rho.reduce((a, b) => a + b, 0);
  // The FP64 solver: the same steps, its state read back in FP64.
  const gpu64 = await bs.GpuEuler2D.create(device, config, { precision: 'fp64' });
  gpu64.setInitialCondition(() => free.state);
  await gpu64.iterate(c.steps);
  out.drag64 = await gpu64.forebodyDrag(free, 0.5);
  const state64 = await gpu64.readState();
  out.state64 = state64.rho instanceof Float64Array ? 'Float64Array' : 'other';
  gpu64.destroy();
  // Implicit steps through the transonic adapter's solver.
  const solve = bs.hipTransonicEngine(device, { sweeps: 2, cfl: 10 });
  const grid = await solve({ config, freeStream: free, body: { nose: { baseRadius: 0.5 } } });
  await grid.start(() => free.state, 1);
  await grid.iterate(c.implicitSteps);
  out.implicitDrag = await grid.drag();
  await grid.checkPrecision();
  grid.destroy();
  try {
    gpu.setInitialCondition(() => ({ rho: 1, u: 0, v: 0, p: -1 }));
    out.nonPhysical = 'accepted';
  } catch (e) {
    out.nonPhysical = e.name;
  }
  gpu.destroy();
  console.log(JSON.stringify(out));
})().catch((e) => {
  console.error(e && e.stack);
  process.exit(1);
});
