'''

Device-resident steps of the HIP engine against the float64 restatement's steps (oracle/), on a
2D solitary wave over a plane beach with the delta-SPH scheme (extrapolated walls, L-corrected
continuity, hydrostatic diffusion, limited viscosity, density floor) and its bed boundary.

Tolerances: the float64 kernels follow the restatement to near rounding; the float32 pair sums
follow it as the reference's float32 device follows its float64 array solver
(tests/testSphGpu.py:503-514: positions 1e-6, velocities 1e-3 over 30 steps). Graph and stream
launches, split and fused steps, and a run alone and in a batch must agree bit for bit.

'''

import math

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


def physicsOf(case) -> oracle.Physics:
    return oracle.Physics(dims = case.dimensions, h = case.smoothingLength, domainMin = case.domainMin,
                          domainMax = case.domainMax, kernelType = case.kernelType, rho0 = case.referenceDensity,
                          c0 = case.soundSpeed, alpha = case.alpha, boundaryFactor = case.boundaryViscosityFactor,
                          gravity = case.gravity, diffusion = case.diffusion, delta = case.delta,
                          extrapolatedWalls = case.extrapolatedWalls, negativePressure = case.negativePressure,
                          correction = case.correction, limiter = case.limiter, densityFloor = case.densityFloor,
                          cfl = case.cfl, maxTimeStep = case.maxTimeStep,
                          hmin = None if case.particles.lengths is None else float(case.particles.lengths.min()))


def smallCase(seed = 0, ppd = 8, **options):
    from pysurf_amd.scenarios import solitaryOnSlope
    return solitaryOnSlope(particlesPerDepth = ppd, seed = seed, **options)


def oracleRun(case, steps):
    p = case.particles
    run = oracle.OracleRun(physicsOf(case), case.boundary.description(), p.positions, p.velocities, p.densities,
                           p.masses, p.isFluid, lengths = p.lengths)
    taken = run.advance(np.inf, steps)
    assert taken == steps
    return run


def engineRun(cases, steps, precision, graph = True, chunk = None, listCapacity = -1):
    from pysurf_amd.batch import BatchEngine
    batch = BatchEngine(cases, precision, listCapacity = listCapacity)
    batch.setClocks(maxSteps = steps)
    batch.advance(chunk or steps, graph)
    if chunk:
        for _ in range(steps // chunk - 1):
            batch.advance(chunk, graph)
    batch.synchronize()
    return batch


@pytest.mark.parametrize('precision, tolerances', [('float64', (1e-10, 1e-8, 1e-9)), ('float32', (1e-6, 1e-3, 1e-3))])
def testDeviceStepsFollowTheOracle(precision, tolerances):
    steps = 60
    print(f'device steps {precision} vs oracle, {steps} steps', flush = True)
    case = smallCase()
    expected = oracleRun(case, steps)
    batch = engineRun([case], steps, precision)
    clock = batch.clocks()
    state = batch.state(0)
    positionTol, velocityTol, densityTol = tolerances
    assert int(clock['steps'][0]) == steps
    assert clock['time'][0] == pytest.approx(expected.time.value, rel = 1e-12 if precision == 'float64' else 1e-5)
    assert state.positions == pytest.approx(expected.positions, abs = positionTol)
    assert state.velocities == pytest.approx(expected.velocities, abs = velocityTol)
    assert state.densities == pytest.approx(expected.dens, abs = densityTol)


def testGraphAndStreamLaunchesAgreeBitForBit():
    case = smallCase(seed = 1)
    print('graph vs stream launches', flush = True)
    a = engineRun([case], 40, 'float32', graph = True, chunk = 8).state(0)
    b = engineRun([case], 40, 'float32', graph = False, chunk = 8).state(0)
    for name in ('positions', 'velocities', 'densities', 'pressures', 'accelerations'):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name


@pytest.mark.parametrize('precision', ['float32', 'float64'])
def testTheNeighborListAndTheCellWalkAgreeBitForBit(precision):
    # A particle whose neighbors overflow the list walks its cells, dealing the same neighbors to the
    # same lanes: a full list, a list of 4 entries (nearly every particle overflows) and no list
    # step identically
    print(f'neighbor list vs cell walk {precision}', flush = True)
    case = smallCase(seed = 8)
    states = [engineRun([case], 25, precision, listCapacity = capacity).state(0) for capacity in (-1, 4, 0)]
    for other in states[1:]:
        for name in ('positions', 'velocities', 'densities', 'pressures'):
            assert np.array_equal(states[0].__dict__[name], other.__dict__[name]), name


@pytest.mark.parametrize('dimensions', [2, 3])
def testTheLdsCopysCapLeavesTheStepAlone(dimensions):
    # A sweep workgroup reads the records it copied into LDS and the rest from HBM, with the same
    # values: the default copy, a 4 KiB one (most neighbors from HBM) and none step identically, in
    # the list build, the three sweeps and the kick the continuity sweep applies to what it reads
    from pysurf_amd.batch import BatchEngine
    print(f'LDS copy cap {dimensions}D', flush = True)
    case = smallCase(seed = 9) if dimensions == 2 else smallRidge(seed = 3, lengthRatio = 1.0)
    states = []
    for cap in (-1, 4096, 0):
        batch = BatchEngine([case], 'float32', lanes = 4 if dimensions == 2 else 8,
                            listCapacity = -1 if dimensions == 2 else 160)
        batch.engine.setNearCells(cap)
        batch.setClocks(maxSteps = 20)
        batch.advance(20)
        batch.synchronize()
        states.append(batch.state(0))
    for other in states[1:]:
        for name in ('positions', 'velocities', 'densities', 'pressures'):
            assert np.array_equal(states[0].__dict__[name], other.__dict__[name]), name
    # a cap beyond the LDS a workgroup may take is refused, naming the range
    with pytest.raises(ValueError, match = '98304'):
        batch.engine.setNearCells(200000)
    with pytest.raises(ValueError, match = '98304'):
        batch.engine.setNearCells(-1, -5)


def testARunAloneAndInABatchAgreeBitForBit():
    # The same run stepped alone and packed between two others of different sizes and physics
    print('batched vs solo', flush = True)
    cases = [smallCase(seed = 2, ppd = 6, waveHeight = 0.3), smallCase(seed = 3, ppd = 8),
             smallCase(seed = 4, ppd = 10, slope = 1.0 / 25.0)]
    steps = 30
    together = engineRun(cases, steps, 'float32')
    for index, case in enumerate(cases):
        alone = engineRun([case], steps, 'float32').state(0)
        mixed = together.state(index)
        for name in ('positions', 'velocities', 'densities'):
            assert np.array_equal(getattr(alone, name), getattr(mixed, name)), (index, name)


def testRunsStopAtTheirOwnClocks():
    from pysurf_amd.batch import BatchEngine
    cases = [smallCase(seed = 5, ppd = 6), smallCase(seed = 6, ppd = 6)]
    batch = BatchEngine(cases, 'float32')
    batch.engine.setClock(0, 0.0, np.inf, 0, 7)
    batch.engine.setClock(1, 0.0, np.inf, 0, 12)
    batch.advance(20)
    clock = batch.clocks()
    assert list(clock['steps']) == [7, 12]


@pytest.mark.parametrize('options', ['delta-SPH', 'antuono-closure', 'filtered'])
def testSplitStepsWithTheHostTimeStepMatchTheDeviceTimeStep(options):
    # residentRates (filtering when the host's step count is due), cflTimeStep on the host and
    # residentIntegrate(dt), against the fused device step that decides both on the device; with the
    # renormalized diffusion and the closure, and with the Shepard filter, as the reference's loop
    # runs them (sphSolver.py:1976-1988)
    from pysurf_amd.pair_sums import PairSums
    case = smallCase(seed = 7)
    steps = 20
    interval = 6 if options == 'filtered' else 0
    print(f'split vs fused steps, {options}', flush = True)

    def make():
        sums = PairSums(2, case.smoothingLength, case.domainMin, case.domainMax, kernelType = case.kernelType)
        diffusion = 'antuono' if options == 'antuono-closure' else case.diffusion
        sums.setDiffusion(diffusion, case.delta, case.soundSpeed, case.referenceDensity, 9.80665)
        sums.setWallPressure(case.extrapolatedWalls)
        sums.setContinuityCorrection(case.correction)
        sums.setViscosityLimiter(case.limiter)
        if options == 'antuono-closure':
            sums.setClosure(True, case.soundSpeed, case.referenceDensity, 1e-6)
        sums.setFilterInterval(interval)
        sums.setPhysics(case.referenceDensity, case.soundSpeed, case.alpha, case.boundaryViscosityFactor, case.gravity,
                        0.0, case.densityFloor)
        particles = case.particles.copy()
        sums.push(particles)
        sums.setBoundary(case.boundary, particles)
        return sums, particles

    split, splitParticles = make()
    time = 0.0
    for step in range(steps):
        vmax, amax = split.residentRates(bool(interval) and step > 0 and step % interval == 0)
        dt = oracle.cflTimeStep(case.smoothingLength, case.soundSpeed, vmax, amax, 0.25, 1.0)
        split.residentIntegrate(dt)
        time += dt
    split.pull(splitParticles)
    fused, fusedParticles = make()
    clock = fused.advance(0.0, np.inf, 0, steps, chunk = steps)
    fused.pull(fusedParticles)
    assert clock['steps'] == steps and clock['time'] == time
    assert np.array_equal(splitParticles.positions, fusedParticles.positions)
    assert np.array_equal(splitParticles.densities, fusedParticles.densities)


@pytest.mark.parametrize('dimensions', [2, 3])
def testBedEnforcementMatchesTheOracle(dimensions):
    # tests/testSphGpu.py:645-680: fluid scattered around a reef profile, a moving piston and both end walls
    from pysurf_amd.bathymetry import Bathymetry
    from pysurf_amd.boundaries import BedBoundary
    from pysurf_amd.pair_sums import PairSums
    from pysurf_amd.particles import ParticleSystem
    rng = np.random.default_rng(7)
    bed = Bathymetry([0.0, 2.0, 5.5, 7.5], [0.0, 0.0, 0.7, 0.7], 1.0)
    s, width = 0.05, (0.6 if dimensions == 3 else None)
    boundary = BedBoundary(bed, s, nLayers = 3, wallTop = 1.5, width = width,
                           endWallTop = 1.5 if dimensions == 2 else None)
    count = 4000
    fluid = np.column_stack([rng.uniform(bed.startX - 0.2, bed.endX + 0.2, count)]
                            + ([rng.uniform(-0.1, 0.7, count)] if dimensions == 3 else [])
                            + [rng.uniform(-1.2, 1.2, count)])
    particles = ParticleSystem.assemble(fluid, boundary.positions(), s, 1000.0, rng.normal(0.0, 1.0, (count, dimensions)))
    boundary.attachPiston(particles)
    boundary.movePiston(particles, bed.startX + 0.03, 0.4)
    moved = particles.copy()
    sums = PairSums(dimensions, 1.3 * s, particles.positions.min(axis = 0), particles.positions.max(axis = 0))
    sums.push(moved)
    sums.setBoundary(boundary, moved)
    sums.enforceBoundary()
    sums.pull(moved)
    physics = oracle.Physics(dims = dimensions, h = 1.3 * s, domainMin = particles.positions.min(axis = 0),
                             domainMax = particles.positions.max(axis = 0))
    x, v = oracle.enforce(physics, boundary.description(pistonSpeed = 0.4), particles.positions, particles.velocities,
                          particles.isFluid)
    assert np.count_nonzero(np.any(x != particles.positions, axis = 1)) > 0.2 * count
    assert moved.positions == pytest.approx(x, abs = 1e-12)
    assert moved.velocities == pytest.approx(v, abs = 1e-12)


@pytest.mark.parametrize('precision, tolerances', [('float64', (1e-12, 1e-13, 1e-11)), ('float32', (1e-6, 1e-7, 1e-5))])
def testGlobalSumsFollowTheOracle(precision, tolerances):
    # The device's float64 sums of its own state against numpy's sums of the same state (the order of
    # summation differs: 1e-12), then the run's energies against the float64 oracle run's after 40 steps
    from oracle import array_formulation as af
    print(f'global sums {precision}', flush = True)
    case = smallCase(seed = 9)
    steps = 40
    batch = engineRun([case], steps, precision)
    sums = batch.globalSums()[0]
    state = batch.state(0)
    masses = case.particles.masses.astype(np.float32 if precision == 'float32' else np.float64)
    g = -case.gravity[-1]
    own = af.globalSums(state.positions, state.velocities, state.densities, masses, state.isFluid, g,
                        case.referenceDensity, case.soundSpeed)
    for k, name in enumerate(('mass', 'kinetic', 'potential', 'elastic')):
        assert sums[k] == pytest.approx(own[name], rel = 1e-12), name
    assert sums[4:6] == pytest.approx(own['momentum'], rel = 1e-12, abs = 1e-12 * own['mass'])
    assert sums[6] == 0.0
    again = batch.globalSums()[0]
    assert np.array_equal(sums, again)
    expected = oracleRun(case, steps)
    reference = af.globalSums(expected.positions, expected.velocities, expected.dens, case.particles.masses,
                              case.particles.isFluid, g, case.referenceDensity, case.soundSpeed)
    kinetic, potential, elastic = tolerances
    print('relative energy differences', sums[1] / reference['kinetic'] - 1, sums[2] / reference['potential'] - 1,
          sums[3] / reference['elastic'] - 1, flush = True)
    assert sums[0] == pytest.approx(reference['mass'], rel = 1e-7)
    assert sums[1] == pytest.approx(reference['kinetic'], rel = kinetic)
    assert sums[2] == pytest.approx(reference['potential'], rel = potential)
    assert sums[3] == pytest.approx(reference['elastic'], rel = elastic)


def testGlobalSumsOfARunAloneAndInABatchAgreeBitForBit():
    cases = [smallCase(seed = 10, ppd = 6), smallCase(seed = 11, ppd = 8)]
    together = engineRun(cases, 10, 'float32').globalSums()
    for index, case in enumerate(cases):
        alone = engineRun([case], 10, 'float32').globalSums()[0]
        assert np.array_equal(alone, together[index]), index


def smallRidge(seed = 0, zone = 0.4, lengthRatio = 0.5, ppd = 8):
    from pysurf_amd.scenarios import ridge3D
    return ridge3D(particlesPerDepth = ppd, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5, ridgeWidth = 0.6,
                   zoneHalfWidth = zone, lengthRatio = lengthRatio, seed = seed)


@pytest.mark.parametrize('precision, tolerances', [('float64', (1e-10, 1e-8, 1e-9)), ('float32', (1e-6, 1e-3, 1e-3))])
@pytest.mark.parametrize('split', ['none', 'half-length'])
def testDeviceStepsFollowTheOracleIn3D(precision, tolerances, split):
    # A solitary wave over a turned Gaussian ridge in a walled tank: the height-field clamp, and with
    # 'half-length' a crest zone split into daughters of half the smoothing length (mixed h)
    steps = 30
    print(f'3D device steps {precision} split {split}, {steps} steps', flush = True)
    case = smallRidge(zone = None if split == 'none' else 0.4)
    assert (case.particles.lengths is None) == (split == 'none')
    expected = oracleRun(case, steps)
    batch = engineRun([case], steps, precision)
    clock = batch.clocks()
    state = batch.state(0)
    positionTol, velocityTol, densityTol = tolerances
    assert int(clock['steps'][0]) == steps
    assert clock['time'][0] == pytest.approx(expected.time.value, rel = 1e-12 if precision == 'float64' else 1e-5)
    assert state.positions == pytest.approx(expected.positions, abs = positionTol)
    assert state.velocities == pytest.approx(expected.velocities, abs = velocityTol)
    assert state.densities == pytest.approx(expected.dens, abs = densityTol)


def testHeightFieldEnforcementMatchesTheOracle():
    from pysurf_amd.bathymetry import HeightField
    from pysurf_amd.boundaries import HeightFieldBoundary
    from pysurf_amd.pair_sums import PairSums
    from pysurf_amd.particles import ParticleSystem
    field = HeightField.sechRidge(0.7, toeX = 1.0, endX = 6.0, halfWidth = 1.5, gridSpacing = 0.1)
    boundary = HeightFieldBoundary(field, 0.05, wallTop = 1.5)
    rng = np.random.default_rng(11)
    count = 5000
    positions = np.column_stack([rng.uniform(-0.2, 6.2, count), rng.uniform(-1.7, 1.7, count),
                                 rng.uniform(-0.2, 1.2, count)])
    velocities = rng.normal(0.0, 1.0, (count, 3))
    fluid = rng.uniform(size = count) < 0.9
    particles = ParticleSystem(positions.copy(), velocities.copy(), np.zeros((count, 3)), np.full(count, 1000.0),
                               np.zeros(count), np.ones(count), fluid)
    sums = PairSums(3, 0.065, positions.min(axis = 0), positions.max(axis = 0))
    sums.push(particles)
    sums.setBoundary(boundary, particles)
    sums.enforceBoundary()
    sums.pull(particles)
    physics = oracle.Physics(dims = 3, h = 0.065, domainMin = positions.min(axis = 0), domainMax = positions.max(axis = 0))
    x, v = oracle.enforce(physics, boundary.description(), positions, velocities, fluid)
    assert np.count_nonzero(np.any(x != positions, axis = 1)) > 0.2 * count
    assert particles.positions == pytest.approx(x, abs = 1e-12)
    assert particles.velocities == pytest.approx(v, abs = 1e-12)


def testSortingTheStateKeepsEveryParticle():
    # The device sort permutes the slots, and the engine hands the particles back in their own order
    from pysurf_amd.batch import BatchEngine
    print('state sort', flush = True)
    cases = [smallCase(seed = 12, ppd = 6), smallCase(seed = 13, ppd = 8)]
    batch = BatchEngine(cases, 'float32')
    batch.setClocks(maxSteps = 10)
    batch.advance(10)
    before = [batch.state(index) for index in range(2)]
    batch.sortParticles()
    for index in range(2):
        after = batch.state(index)
        for name in ('positions', 'velocities', 'densities'):
            assert np.array_equal(getattr(before[index], name), getattr(after, name)), (index, name)
    identity = batch.identity.cpu().numpy()
    assert not np.array_equal(identity, np.arange(batch.count))
    assert np.array_equal(np.sort(identity[:cases[0].count]), np.arange(cases[0].count))


@pytest.mark.parametrize('precision, tolerances', [('float64', (1e-10, 1e-8, 1e-9)), ('float32', (1e-6, 1e-3, 1e-3))])
def testSortedStepsFollowTheOracle(precision, tolerances):
    # Sorting every 10 steps changes only the order of the sums' terms
    from pysurf_amd.batch import BatchEngine
    print(f'sorted steps {precision}', flush = True)
    case = smallCase(seed = 14)
    steps = 60
    expected = oracleRun(case, steps)
    batch = BatchEngine([case], precision)
    batch.setClocks(maxSteps = steps)
    for _ in range(steps // 10):
        batch.advance(10)
        batch.sortParticles()
    state = batch.state(0)
    positionTol, velocityTol, densityTol = tolerances
    assert int(batch.clocks()['steps'][0]) == steps
    assert state.positions == pytest.approx(expected.positions, abs = positionTol)
    assert state.velocities == pytest.approx(expected.velocities, abs = velocityTol)
    assert state.densities == pytest.approx(expected.dens, abs = densityTol)


def testSortedRunsAloneAndInABatchAgreeBitForBit():
    from pysurf_amd.batch import BatchEngine
    cases = [smallCase(seed = 15, ppd = 6), smallCase(seed = 16, ppd = 8)]

    def stepped(selection):
        batch = BatchEngine(selection, 'float32')
        batch.setClocks(maxSteps = 30)
        for _ in range(3):
            batch.advance(10)
            batch.sortParticles()
        return [batch.state(index) for index in range(len(selection))]

    together = stepped(cases)
    for index, case in enumerate(cases):
        alone = stepped([case])[0]
        for name in ('positions', 'velocities', 'densities'):
            assert np.array_equal(getattr(alone, name), getattr(together[index], name)), (index, name)
