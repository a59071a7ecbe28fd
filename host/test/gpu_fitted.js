'use strict';
// Runs a fitted-shock case through the Node facade on the GPU and prints
// the results as JSON: `node gpu_fitted.js case.json`. The case holds the
// lines, their fractions, the starting heights and the free stream; the
// calling test compares the heights and the shock with the Python facade
// on the same case.
const fs = require('fs');
const bs = require('..');

(async () => {
  const c = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  const free = bs.freeStreamAt(c.mach, 1.4);
  const ni = c.lines.length - 1;
  const nj = c.nj;
  const lines = c.lines.map((l, i) => ({ x: l[0], y: l[1], dx: l[2], dy: l[3], fractions: c.fractions[i] }));
  const node = (i, j) => {
    const l = lines[i];
    const s = c.heights[i] * l.fractions[j];
    return [l.x + s * l.dx, l.y + s * l.dy];
  };
  const x = new Float64Array((ni + 1) * (nj + 1));
  const y = new Float64Array((ni + 1) * (nj + 1));
  for (let j = 0; j <= nj; j++) {
    for (let i = 0; i <= ni; i++) {
      const [a, b] = node(i, j);
      x[i + j * (ni + 1)] = a;
      y[i + j * (ni + 1)] = b;
    }
  }
  const states = Array.from({ length: ni }, () => free.state);
  const config = {
    grid: { ni, nj, x, y },
    geometry: 'axisymmetric',
    gamma: 1.4,
    flux: 'hllc',
    limiter: 'vanAlbada',
    reconstruction: 'sharedScaled',
    timeIntegrator: 'ssprk2',
    cfl: 0.5,
    timeStepping: 'local',
    boundaries: {
      iMin: { kind: 'inflow', state: free.state },
      iMax: { kind: 'outflow' },
      jMin: { kind: 'wall' },
      jMax: { kind: 'faceStates', states },
    },
    fittedShock: { freeStream: free.state },
  };
  const device = new bs.HipDevice(0);
  const gpu = await bs.GpuEuler2D.create(device, config, { precision: 'fp64' });
  const lineGrid = await bs.GpuLineGrid.create(device, lines, nj, {
    target: gpu.geometryBuffer,
    boundary: gpu.boundaryBuffer,
    firstNode: 'fixed',
  });
  lineGrid.setHeights(c.heights);
  gpu.setInitialCondition(() => free.state);
  await gpu.iterateFitted(lineGrid, c.moves, 1);
  const out = { heights: await lineGrid.readHeights() };
  const shock = await gpu.readShock();
  out.speed = Array.from(shock.speed);
  const state = await gpu.readState();
  out.steps = state.steps;
  try {
    await gpu.iterateImplicit(1, { cfl: 10, sweeps: 2 });
    out.implicit = 'allowed';
  } catch (e) {
    out.implicit = e.message;
  }
  gpu.destroy();
  console.log(JSON.stringify(out));
})().catch((e) => {
  console.error(e && e.stack);
  process.exit(1);
});
