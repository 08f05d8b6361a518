'''

The benchmark configurations at their own sizes on the device, beside the small cases the other
GPU tests step:

  the batched sweep   all 256 runs (5.07 M particles) in one engine at the autotuner's launch shape
                      for this workload (4 lanes, 256 threads): a 16 and a 48 particles-per-depth
                      member follow the float64 restatement stepped from the same state, one member
                      steps bit for bit alone as in the mix, and every run keeps its mass
  the 3D crest zone   daughters that keep the mother's smoothing length (length ratio 1), so a
                      crest particle has more neighbors than the list holds and walks its cells:
                      the steps follow the restatement
  the 3D ridge        all 12.4 M particles of the 32 particles-per-depth ridge with its split crest
                      zone (daughters keep h, up to about 600 neighbors, more than the list's 768
                      slots hold at the densest), stepped at the bench's launch shape: the state stays
                      finite, the mass is the start's, the energy and the momentum along the wave
                      move by no more than the start's relaxation allows, and the fluid stays in the
                      tank
  the request loop    a full-size 2D sweep run and a 16 particles-per-depth 3D ridge run of 0.73 M
                      particles under an HBM budget that holds one of them at a time, so the second
                      queues: both are served, and each history is the run's alone, bit for bit

'''

import numpy as np
import pytest

from test_gpu_steps import oracleRun

pytestmark = pytest.mark.gpu

LANES, BLOCK = 4, 256


def stepped(cases, steps: int, precision: str = 'float32'):
    from pysurf_amd.batch import BatchEngine
    batch = BatchEngine(cases, precision, blockSize = BLOCK, lanes = LANES)
    batch.setClocks(maxSteps = steps)
    batch.advance(steps)
    batch.synchronize()
    return batch


@pytest.mark.timeout(600)
def testTheBatchedSweepOf256RunsFollowsTheOracle():
    from pysurf_amd.scenarios import sweepRun
    cases = [sweepRun(seed) for seed in range(256)]
    steps = 20
    print(f'batched sweep: {len(cases)} runs, {sum(c.count for c in cases)} particles, {steps} steps', flush = True)
    batch = stepped(cases, steps)
    clock = batch.clocks()
    assert np.all(clock['steps'] == steps)
    sums = batch.globalSums()
    masses = np.array([case.particles.masses[case.particles.isFluid].sum() for case in cases])
    assert sums[:, 0] == pytest.approx(masses, rel = 1e-6)
    assert np.all(np.isfinite(sums))
    resolutions = [case.parameters.get('particlesPerDepth') for case in cases]
    for ppd in (16, 48):
        index = resolutions.index(ppd)
        case = cases[index]
        print(f'member {index} ({ppd} ppd, {case.count} particles) against the oracle', flush = True)
        expected = oracleRun(case, steps)
        state = batch.state(index)
        assert clock['time'][index] == pytest.approx(expected.time.value, rel = 1e-5)
        assert state.positions == pytest.approx(expected.positions, abs = 1e-6)
        assert state.velocities == pytest.approx(expected.velocities, abs = 1e-3)
        assert state.densities == pytest.approx(expected.dens, abs = 1e-3)
    index = resolutions.index(48)
    print(f'member {index} alone', flush = True)
    alone = stepped([cases[index]], steps).state(0)
    mixed = batch.state(index)
    for name in ('positions', 'velocities', 'densities', 'pressures'):
        assert np.array_equal(getattr(alone, name), getattr(mixed, name)), name


@pytest.mark.timeout(600)
def testTheFullSizeRidgeStaysPhysical():
    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.scenarios import ridge3D
    case = ridge3D()
    steps = 20
    print(f'3D ridge at full size: {case.count} particles, {steps} steps', flush = True)
    batch = BatchEngine([case], 'float32', blockSize = 256, lanes = 8, listCapacity = 768)
    before = batch.globalSums()[0]
    batch.setClocks(maxSteps = steps)
    batch.advance(steps)
    batch.synchronize()
    clock = batch.clocks()
    after = batch.globalSums()[0]
    assert int(clock['steps'][0]) == steps and clock['time'][0] > 0.0
    assert clock['failed'][0] == 0
    assert np.all(np.isfinite(after))
    # mass is carried by the particles, each keeping its own
    assert after[0] == pytest.approx(before[0], rel = 1e-12)
    # kinetic + potential + elastic: over the first 20 steps the start's pressure field relaxes
    # (elastic energy turns kinetic) and the delta-SPH diffusion dissipates a little; the total moves
    # by about 2e-4 of itself
    energy = lambda sums: sums[1] + sums[2] + sums[3]
    assert abs(energy(after) - energy(before)) < 1e-3 * abs(energy(before))
    # the wave's momentum along x moves only by what the bed and walls push in a few steps
    assert abs(after[4] - before[4]) < 0.02 * abs(before[4])
    state = batch.state(0)
    fluid = state.isFluid
    assert np.all(np.isfinite(state.positions)) and np.all(np.isfinite(state.velocities))
    low, high = np.asarray(case.domainMin), np.asarray(case.domainMax)
    margin = 2.0 * case.smoothingLength
    assert np.all(state.positions[fluid] > low - margin) and np.all(state.positions[fluid] < high + margin)
    assert state.densities[fluid].min() >= case.densityFloor * (1.0 - 1e-12)


@pytest.mark.timeout(300)
@pytest.mark.parametrize('precision, tolerances', [('float64', (1e-10, 1e-8, 1e-9)), ('float32', (1e-6, 1e-3, 1e-3))])
def testACrestZoneWhoseDaughtersKeepTheirLengthFollowsTheOracle(precision, tolerances):
    from oracle import oracle
    from pysurf_amd.batch import BatchEngine
    from test_gpu_steps import physicsOf, smallRidge
    case = smallRidge(zone = 0.4, lengthRatio = 1.0)
    capacity = 160
    offsets, _ = oracle.neighbors(physicsOf(case), case.particles.positions, case.particles.lengths)
    counts = np.diff(offsets)
    print(f'3D crest zone, length ratio 1, {precision}: {case.count} particles, neighbors up to {counts.max()} '
          f'({np.count_nonzero(counts > capacity)} over the list of {capacity})', flush = True)
    assert np.count_nonzero(counts > capacity) > 100
    steps = 20
    expected = oracleRun(case, steps)
    batch = BatchEngine([case], precision, listCapacity = capacity)
    batch.setClocks(maxSteps = steps)
    batch.advance(steps)
    batch.synchronize()
    state = batch.state(0)
    positionTol, velocityTol, densityTol = tolerances
    assert int(batch.clocks()['steps'][0]) == steps
    assert state.positions == pytest.approx(expected.positions, abs = positionTol)
    assert state.velocities == pytest.approx(expected.velocities, abs = velocityTol)
    assert state.densities == pytest.approx(expected.dens, abs = densityTol)


@pytest.mark.timeout(600)
def testARequestThatQueuesIsServedAsItWouldBeAlone():
    from pysurf_amd.scenarios import servingRidge, sweepRun
    from pysurf_amd.serving import Request, Server, runToBreaking
    ridge, sweep = servingRidge(4), sweepRun(44)
    ends = {'ridge': 0.05, 'sweep': 0.1}
    server = Server()
    # room for the ridge or the sweep run, not both: the sweep run, arriving second, waits
    server.budgetBytes = server.footprint(ridge) + 0.5 * server.footprint(sweep)
    print(f'serving: ridge {ridge.count} particles, sweep {sweep.count}, budget {server.budgetBytes / 1e9:.2f} GB',
          flush = True)
    requests = [Request(key = 'ridge', arrival = 0.0, case = ridge, endTime = ends['ridge']),
                Request(key = 'sweep', arrival = 0.0, case = sweep, endTime = ends['sweep'])]
    served = {record.key: record for record in server.serve(requests, wallLimit = 300.0)}
    assert sorted(served) == ['ridge', 'sweep']
    assert all(record.refused == '' for record in served.values())
    assert max(depth for _, depth in server.queueDepth) == 1
    assert served['sweep'].admitted >= served['ridge'].finished
    for key, case in (('ridge', ridge), ('sweep', sweep)):
        capacity = server.listCapacity2D if case.dimensions == 2 else server.listCapacity3D
        alone = runToBreaking([case], 'float32', endTime = ends[key], lanes = server.lanes, listCapacity = capacity)[0]
        mixed = served[key].tracker
        print(f'{key}: {len(alone.times)} readings alone, {len(mixed.times)} served', flush = True)
        assert len(alone.times) > 5
        for name in ('times', 'crestX', 'crestHeight', 'lean', 'faceSlope', 'waterline'):
            assert np.array_equal(np.asarray(getattr(alone, name)), np.asarray(getattr(mixed, name)), equal_nan = True), \
                (key, name)
