'use strict';
// Host-side logic of the Node facade, runnable without a GPU: the addon
// loads and exports its functions, errors map to the reference's classes,
// the exclusive runner refuses overlapping calls, configurations are
// checked before anything reaches the device, and the helpers restate
// verified-drag.ts / forces.ts.
const assert = require('assert');
const bs = require('..');

const results = [];
async function test(name, fn) {
  try {
    await fn();
    results.push(['ok', name]);
  } catch (e) {
    results.push(['FAIL', name, e && e.stack]);
  }
}

(async () => {
  await test('addon exports every binding', () => {
    const names = ['version', 'ctxCreate', 'ctxDestroy', 'solverCreate', 'solverDestroy', 'computeMetrics',
      'setPrimitives', 'setState', 'readState', 'setFaceStates', 'setFaceFluxes', 'residualNorm',
      'residualVector', 'wallFluxes', 'forebodyDrag', 'setImplicitCfl', 'iterate', 'advanceTo', 'iterateImplicit',
      'readState64', 'iterateFitted', 'setFittedShock', 'readShock', 'setLineGrid', 'setHeights', 'moveGrid',
      'readHeights', 'batchCreate', 'batchDestroy', 'batchSetUniform', 'batchSetPrimitives', 'batchReadState',
      'batchSetActive', 'batchSetCoarse', 'batchSetInt', 'batchSetImplicitCfl', 'batchIterate', 'batchIterateImplicit',
      'batchResidualNorms', 'batchDrags', 'batchStatus', 'checkPrecision'];
    const a = bs.native();
    for (const n of names) assert.strictEqual(typeof a[n], 'function', n);
    assert.ok(/gfx950/.test(a.version()));
  });

  await test('error codes map to the reference classes', () => {
    const mk = (code) => Object.assign(new Error('m'), { code });
    assert.ok(bs.translate(mk('BS_ERR_RANGE')) instanceof RangeError);
    assert.ok(bs.translate(mk('BS_ERR_NONPHYSICAL')) instanceof bs.NonPhysicalStateError);
    assert.ok(bs.translate(mk('BS_ERR_MEMORY')) instanceof bs.GpuMemoryLimitError);
    assert.ok(bs.translate(mk('BS_ERR_EXECUTION')) instanceof bs.GpuExecutionError);
    assert.ok(bs.translate(mk('BS_ERR_PRECISION')) instanceof bs.GpuPrecisionError);
    assert.strictEqual(bs.translate(mk('BS_ERR_NONPHYSICAL')).name, 'NonPhysicalStateError');
  });

  await test('the exclusive runner refuses overlapping operations', async () => {
    const r = new bs.ExclusiveRunner('GpuEuler2D');
    let release;
    const first = r.run(() => new Promise((resolve) => { release = resolve; }));
    await assert.rejects(r.run(async () => 1), /await the previous operation first/);
    assert.throws(() => r.assertIdle(), /await the previous operation first/);
    release(7);
    assert.strictEqual(await first, 7);
    assert.strictEqual(await r.run(async () => 2), 2);
  });

  const grid = { ni: 2, nj: 2, x: [0, 1, 2, 0, 1, 2, 0, 1, 2], y: [0, 0, 0, 1, 1, 1, 2, 2, 2] };
  const free = { rho: 1, u: 2, v: 0, p: 1 };
  const base = {
    grid, geometry: 'axisymmetric', gamma: 1.4, flux: 'hllc', limiter: 'minmod', reconstruction: 'perVariable',
    timeIntegrator: 'ssprk3', cfl: 0.5, timeStepping: 'local',
    boundaries: { iMin: { kind: 'inflow', state: free }, iMax: { kind: 'outflow' }, jMin: { kind: 'wall' },
      jMax: { kind: 'inflow', state: free } },
  };

  await test('configurations are checked on the host', () => {
    const c = bs.nativeConfig(base);
    assert.strictEqual(c.limiter, 1);
    assert.strictEqual(c.boundaries[0].kind, 0);
    assert.strictEqual(c.boundaries[2].kind, 2);
    assert.strictEqual(c.x.length, 9);
    assert.throws(() => bs.nativeConfig(Object.assign({}, base, { flux: 'godunov' })), RangeError);
    assert.throws(() => bs.nativeConfig(Object.assign({}, base, { limiter: 'superbee' })), RangeError);
    assert.throws(() => bs.nativeConfig(Object.assign({}, base, { source: new Float64Array(16) })), RangeError);
    // A fitted shock is set on the solver after it is made, not in the plain data.
    assert.strictEqual(bs.nativeConfig(Object.assign({}, base, { fittedShock: { freeStream: free } })).ni, 2);
  });

  await test('a line grid is made on a HIP solver only', async () => {
    await assert.rejects(() => bs.GpuLineGrid.create(0, [], 2, { target: {} }), RangeError);
  });

  await test('metrics come from the library without a GPU', () => {
    const m = bs.computeMetrics(grid, 'axisymmetric');
    assert.strictEqual(m.iFaces.nx.length, 6);
    assert.strictEqual(m.volume[0], 0.5);  // unit cell, centroid radius 0.5
    assert.strictEqual(m.jFaces.ny[0], 1);
    assert.throws(() => bs.computeMetrics({ ni: 2, nj: 2, x: grid.x, y: grid.y.map((v) => v - 1) }, 'axisymmetric'),
      RangeError);
  });

  await test('freeStreamAt and forebodyPressureDrag restate the reference', () => {
    const f = bs.freeStreamAt(3, 1.4);
    assert.strictEqual(f.state.u, Math.fround(3 * Math.sqrt(1.4)));
    assert.strictEqual(f.dynamicPressure, 0.5 * f.state.u * f.state.u);
    const wall = [{ x: 0, y: 0, nx: 1, ny: 0, area: 0.1, flux: [0, 2, 0, 0] },
      { x: 0, y: 0, nx: 0.5, ny: 0.5, area: 0.2, flux: [0, 0.5, 0, 0] }];
    assert.ok(Math.abs(bs.forebodyPressureDrag(wall, { pressure: 0, dynamicPressure: 1 }, 1) - 0.6) < 1e-15);
  });

  await test('solveBatch applies the stopping rule case by case', async () => {
    // A stand-in batch with scripted readings: case 0 reaches the residual
    // target 1e-6 with a settled drag, case 1 stalls at a floor
    // with a settled drag, case 2 never settles, case 3 turns non-physical.
    const checks = [];
    const fake = {
      count: 4,
      done: 0,
      active: null,
      setActive(a) { this.active = a.slice(); },
      async iterate(n) { this.done += n; checks.push(this.active.slice()); },
      async status() {
        return { status: Uint32Array.from([0, 0, 0, this.done >= 240 ? 1 : 0]), steps: Float64Array.from([1, 1, 1, 1], () => this.done) };
      },
      async densityResidualNorms() {
        const k = this.done / 120;
        return Float64Array.from([10 ** (-2 * k), 1e-3, 1 / k, 1]);
      },
      async forebodyDrags() {
        const k = this.done / 120;
        return Float64Array.from([0.2, 0.3, 0.2 + 0.01 * k, 0.4]);
      },
    };
    const res = await bs.solveBatch(fake, { pressure: 1, dynamicPressure: 1 }, 0.5, { chunk: 120, maxSteps: 1200 });
    assert.deepStrictEqual(res.map((r) => r.reason), ['converged', 'converged (residual floor)', 'step limit', 'non-physical state']);
    assert.strictEqual(res[0].steps, 360);      // 1e-6 at the 3rd check, the drag settled over 2
    assert.strictEqual(res[1].steps, 600);      // no new minimum in 4 checks after the first
    assert.strictEqual(res[2].steps, 1200);
    assert.strictEqual(res[3].steps, 240);
    // Finished cases leave the launches at the next check.
    assert.deepStrictEqual(checks[checks.length - 1], [false, false, true, false]);
  });

  await test('a device that does not exist is an error, not a fallback', () => {
    let threw = false;
    try {
      new bs.HipDevice(99);
    } catch (e) {
      threw = e instanceof RangeError || e instanceof bs.GpuExecutionError;
    }
    assert.ok(threw);
  });

  await test('destroy is refused while an operation is pending', async () => {
    // Engines made without a device: the guard runs before any native call.
    for (const Cls of [bs.GpuEuler2D, bs.GpuBatch2D]) {
      const e = Object.create(Cls.prototype);
      e.exclusive = new bs.ExclusiveRunner(Cls.name);
      e.handle = null;
      let release;
      const pending = e.exclusive.run(() => new Promise((resolve) => { release = resolve; }));
      assert.throws(() => e.destroy(), /await the previous operation first/);
      release();
      await pending;
      e.destroy();
      assert.strictEqual(e.handle, null);
    }
  });

  await test('heights and activity masks of the wrong length are RangeErrors', () => {
    const lg = Object.create(bs.GpuLineGrid.prototype);
    lg.ni = 4;
    assert.throws(() => lg.heightsOf([1, 1, 1, 1]), RangeError);
    assert.strictEqual(lg.heightsOf([1, 1, 1, 1, 1]).length, 5);
    const b = Object.create(bs.GpuBatch2D.prototype);
    b.exclusive = new bs.ExclusiveRunner('GpuBatch2D');
    b.count = 3;
    b.handle = null;
    assert.throws(() => b.setActive([true, false]), RangeError);
  });

  for (const r of results) console.log(r[0], r[1], r[2] || '');
  process.exit(results.every((r) => r[0] === 'ok') ? 0 : 1);
})();
