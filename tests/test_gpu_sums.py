'''

The HIP engine's host-step sums against the float64 restatement (oracle/), on the reference's
test clouds (tests/testSphGpu.py:32-62: 600 or 900 particles 25 m from the origin, 80% fluid), in
2D, 3D and periodic 3D, with both kernels, in both pair precisions. The float32 tolerance is the
reference's own, 1e-5 of the largest expected value (tests/testSphGpu.py:101); the float64 kernels
are held to 1e-10. Neighbor sets are compared bit for bit with the restatement of each precision's
predicate.

'''

import numpy as np
import pytest

from clouds import alpha, c0, cloud, expectedShifts, physicsFor, rho0, shiftingBlock, strainedFlow
from oracle import oracle

pytestmark = pytest.mark.gpu

cases = [(2, False), (3, False), (3, True)]
kernelTypes = ['cubicSpline', 'wendlandC2']
precisions = ['float32', 'float64']
tolerance = {'float32': 1e-5, 'float64': 1e-10}


def pairSums(case, precision, lengths = None):
    from pysurf_amd.pair_sums import PairSums
    positions = case['positions']
    sums = PairSums(positions.shape[1], case['h'], positions.min(axis = 0), positions.max(axis = 0),
                    case['period'] or 0.0, 'float64', 'cuda:0', case['kernelType'], pairPrecision = precision)
    sums.update(positions, case['masses'], case['isFluid'], lengths)
    return sums


def gravityOf(dimensions):
    g = np.zeros(dimensions)
    g[-1] = -9.81
    return g


@pytest.mark.parametrize('precision', precisions)
@pytest.mark.parametrize('dimensions, periodic', cases)
@pytest.mark.parametrize('kernelType', kernelTypes)
def testAccelerationsMatchTheOracle(precision, dimensions, periodic, kernelType):
    print(f'accelerations {precision} {dimensions}D periodic={periodic} {kernelType}', flush = True)
    case = cloud(dimensions, periodic, kernelType = kernelType)
    g = gravityOf(dimensions)
    physics = physicsFor(case, boundaryFactor = 0.3)
    expected = oracle.accelerations(physics, case['positions'], case['velocities'], case['densities'],
                                    case['pressures'], case['masses'], case['isFluid'])
    result = pairSums(case, precision).pairAccelerations(case['velocities'], case['densities'], case['pressures'], c0,
                                                         alpha, 0.3, g)
    assert result == pytest.approx(expected, abs = tolerance[precision] * np.abs(expected).max())


@pytest.mark.parametrize('precision', precisions)
@pytest.mark.parametrize('dimensions, periodic', cases)
@pytest.mark.parametrize('kernelType', kernelTypes)
def testLimitedAccelerationsMatchTheOracle(precision, dimensions, periodic, kernelType):
    print(f'limited accelerations {precision} {dimensions}D periodic={periodic} {kernelType}', flush = True)
    case = cloud(dimensions, periodic, kernelType = kernelType)
    velocities = strainedFlow(case)
    g = gravityOf(dimensions)
    physics = physicsFor(case, boundaryFactor = 0.3, limiter = True)
    gradients = oracle.velocityGradients(physics, case['positions'], velocities, case['densities'], case['masses'],
                                         case['isFluid'])
    expected = oracle.accelerations(physics, case['positions'], velocities, case['densities'], case['pressures'],
                                    case['masses'], case['isFluid'], gradients = gradients)
    sums = pairSums(case, precision)
    sums.setViscosityLimiter(True)
    matrices = sums.velocityGradients(velocities, case['densities'])
    result = sums.pairAccelerations(velocities, case['densities'], case['pressures'], c0, alpha, 0.3, g)
    assert matrices == pytest.approx(gradients, abs = 2 * tolerance[precision] * np.abs(gradients).max())
    assert result == pytest.approx(expected, abs = tolerance[precision] * np.abs(expected).max())


@pytest.mark.parametrize('precision', precisions)
@pytest.mark.parametrize('dimensions, periodic', cases)
@pytest.mark.parametrize('kernelType', kernelTypes)
@pytest.mark.parametrize('corrected', [False, True])
@pytest.mark.parametrize('mode', ['none', 'molteni', 'hydrostatic'])
def testContinuityMatchesTheOracle(precision, dimensions, periodic, kernelType, corrected, mode):
    print(f'continuity {precision} {dimensions}D periodic={periodic} {kernelType} L={corrected} {mode}', flush = True)
    case = cloud(dimensions, periodic, kernelType = kernelType)
    physics = physicsFor(case, correction = corrected, diffusion = mode, delta = 0.1)
    corrections = None
    if corrected:
        corrections = oracle.correctionMatrices(physics, case['positions'], case['densities'], case['masses'],
                                                case['isFluid'])
    expected = oracle.continuityRate(physics, case['positions'], case['velocities'], case['densities'], case['masses'],
                                     case['isFluid'], corrections = corrections)
    sums = pairSums(case, precision)
    sums.setDiffusion(mode, 0.1, c0, rho0, 9.81)
    sums.setContinuityCorrection(corrected)
    if corrected:
        matrices = sums.correctionMatrices(case['densities'])
        assert matrices == pytest.approx(corrections, abs = 2 * tolerance[precision] * np.abs(corrections).max())
    result = sums.continuityRate(case['velocities'], case['densities'])
    assert result == pytest.approx(expected, abs = tolerance[precision] * np.abs(expected).max())


@pytest.mark.parametrize('precision', precisions)
@pytest.mark.parametrize('dimensions, periodic', cases)
@pytest.mark.parametrize('kernelType', kernelTypes)
@pytest.mark.parametrize('corrected', [False, True])
def testRenormalizedDiffusionMatchesTheOracle(precision, dimensions, periodic, kernelType, corrected):
    # tests/testSphGpu.py:184-205: with the correction off the device builds L for the gradients
    # alone; the gradients to 2e-4 of the largest in float32, the rate to 1e-5
    print(f'antuono {precision} {dimensions}D periodic={periodic} {kernelType} L={corrected}', flush = True)
    case = cloud(dimensions, periodic, kernelType = kernelType)
    fluid = case['isFluid']
    physics = physicsFor(case, correction = corrected, diffusion = 'antuono', delta = 0.1)
    L = oracle.correctionMatrices(physics, case['positions'], case['densities'], case['masses'], fluid)
    gradients = oracle.densityGradients(physics, case['positions'], case['densities'], case['masses'], fluid, L)
    expected = oracle.continuityRate(physics, case['positions'], case['velocities'], case['densities'], case['masses'],
                                     fluid, corrections = L if corrected else None, densityGradients = gradients)
    sums = pairSums(case, precision)
    sums.setDiffusion('antuono', 0.1, c0, rho0, 9.81)
    sums.setContinuityCorrection(corrected)
    result = sums.continuityRate(case['velocities'], case['densities'])
    deviceGradients = sums.densityGradients(case['densities'])
    gradientTolerance = 2e-4 if precision == 'float32' else 1e-10
    assert deviceGradients == pytest.approx(gradients, abs = gradientTolerance * np.abs(gradients).max())
    assert np.all(deviceGradients[~fluid] == 0.0)
    assert result == pytest.approx(expected, abs = tolerance[precision] * np.abs(expected).max())
    # The correction matrices serve the gradients with the continuity correction off
    assert sums.correctionMatrices(case['densities']) == pytest.approx(L, abs = 2 * tolerance[precision])


@pytest.mark.parametrize('precision', precisions)
@pytest.mark.parametrize('dimensions, periodic', cases)
@pytest.mark.parametrize('mode', ['hydrostatic', 'antuono'])
def testClosureMatchesTheOracle(precision, dimensions, periodic, mode):
    # tests/testSphGpu.py:216-290: the coefficients from G of a strained flow, then the accelerations
    # with the closure's viscosity and the continuity with each pair's harmonic-mean delta
    print(f'closure {precision} {dimensions}D periodic={periodic} {mode}', flush = True)
    case = cloud(dimensions, periodic)
    fluid = case['isFluid']
    velocities = strainedFlow(case)
    nu = 1.0034e-6
    physics = physicsFor(case, boundaryFactor = 0.3, diffusion = mode, delta = 0.1)
    G = oracle.velocityGradients(physics, case['positions'], velocities, case['densities'], case['masses'], fluid)
    alphas, deltas = oracle.closureCoefficients(physics, G, fluid, c0)
    closure = dict(alphas = alphas, deltas = deltas, soundSpeed = c0, referenceDensity = rho0, kinematicViscosity = nu)
    expected = oracle.accelerations(physics, case['positions'], velocities, case['densities'], case['pressures'],
                                    case['masses'], fluid, closure = closure)
    densityGradients = None
    if mode == 'antuono':
        L = oracle.correctionMatrices(physics, case['positions'], case['densities'], case['masses'], fluid)
        densityGradients = oracle.densityGradients(physics, case['positions'], case['densities'], case['masses'], fluid,
                                                   L)
    expectedRate = oracle.continuityRate(physics, case['positions'], velocities, case['densities'], case['masses'],
                                         fluid, densityGradients = densityGradients, closure = closure)
    sums = pairSums(case, precision)
    sums.setDiffusion(mode, 0.1, c0, rho0, 9.81)
    sums.setClosure(True, c0, rho0, nu)
    deviceAlphas, deviceDeltas = sums.closureCoefficients(velocities, case['densities'])
    scale = tolerance[precision]
    assert deviceAlphas == pytest.approx(alphas, abs = 20 * scale * 0.2)
    assert deviceDeltas == pytest.approx(deltas, abs = 20 * scale * 0.2)
    cap = float(np.float32(0.2)) if precision == 'float32' else 0.2
    assert np.all(deviceAlphas[~fluid] == 0.0) and np.any(deviceDeltas[fluid] == cap)
    result = sums.pairAccelerations(velocities, case['densities'], case['pressures'], c0, alpha, 0.3,
                                    gravityOf(dimensions))
    assert result == pytest.approx(expected, abs = scale * np.abs(expected).max())
    rate = sums.continuityRate(velocities, case['densities'])
    assert rate == pytest.approx(expectedRate, abs = scale * np.abs(expectedRate).max())


@pytest.mark.parametrize('precision', precisions)
@pytest.mark.parametrize('dimensions', [2, 3])
@pytest.mark.parametrize('kernelType', kernelTypes)
@pytest.mark.parametrize('local', [False, True])
def testShiftsMatchTheOracle(precision, dimensions, kernelType, local):
    # tests/testSphGpu.py:318-359: the shifts to 1e-4 of the largest, the density changes to 1e-3 and
    # the velocity changes to 1e-4 in float32 (the reference's own bounds); 1e-10 in float64
    print(f'shifting {precision} {dimensions}D {kernelType} local={local}', flush = True)
    block = shiftingBlock(dimensions, kernelType)
    fluid = block['isFluid']
    size = 2.0 * block['h'] * block['timeStep'] if local else block['size']
    physics = physicsFor(block)
    shift, change, velocity = oracle.shiftingDisplacement(physics, block['positions'], block['velocities'],
                                                          block['densities'], block['masses'], fluid, size, local,
                                                          block['spacing'])
    surface = expectedShifts(block)[3]
    assert (fluid & ~surface & (np.abs(shift).sum(axis = 1) > 0)).sum() > 50
    sums = pairSums(block, precision)
    sums.setDiffusion('none', 0.1, c0, rho0, 9.81)
    sums.setShifting(True, block['mach'], block['spacing'], c0, local)
    result, densityChange, velocityChange = sums.shiftingDisplacement(block['densities'], block['timeStep'],
                                                                      block['velocities'])
    loose = precision == 'float32'
    assert result == pytest.approx(shift, abs = (1e-4 if loose else 1e-10) * np.abs(shift).max())
    assert densityChange == pytest.approx(change, abs = (1e-3 if loose else 1e-10) * np.abs(change).max())
    assert velocityChange == pytest.approx(velocity, abs = (1e-4 if loose else 1e-10) * np.abs(velocity).max())
    assert np.all(result[surface | ~fluid] == 0.0)
    sums.setPhysics(rho0, c0, alpha, 0.0, gravityOf(dimensions), 0.0, 0.0)
    with pytest.raises(ValueError, match = 'shift'):
        sums.residentRates()


@pytest.mark.parametrize('precision', precisions)
@pytest.mark.parametrize('dimensions, periodic', cases)
@pytest.mark.parametrize('kernelType', kernelTypes)
def testWallPressureMatchesTheOracle(precision, dimensions, periodic, kernelType):
    print(f'wall pressure {precision} {dimensions}D periodic={periodic} {kernelType}', flush = True)
    case = cloud(dimensions, periodic, kernelType = kernelType)
    g = gravityOf(dimensions)
    expected, density = oracle.wallPressure(physicsFor(case), case['positions'], case['densities'], case['pressures'],
                                            case['masses'], case['isFluid'])
    sums = pairSums(case, precision)
    result = sums.wallPressure(case['densities'], case['pressures'], g, rho0, c0)
    assert result == pytest.approx(expected, abs = tolerance[precision] * np.abs(expected).max())
    boundary = ~case['isFluid']
    assert sums.wallDensities[boundary] == pytest.approx(density[boundary], rel = tolerance[precision])


@pytest.mark.parametrize('precision', precisions)
@pytest.mark.parametrize('dimensions, periodic', cases)
@pytest.mark.parametrize('kernelType', kernelTypes)
def testShepardMatchesTheOracle(precision, dimensions, periodic, kernelType):
    # tests/testSphGpu.py:377-386: two float32 sums over up to 270 neighbors, relative 2e-6
    print(f'shepard {precision} {dimensions}D periodic={periodic} {kernelType}', flush = True)
    case = cloud(dimensions, periodic, kernelType = kernelType)
    expected = oracle.shepardDensity(physicsFor(case), case['positions'], case['densities'], case['masses'],
                                     case['isFluid'])
    result = pairSums(case, precision).shepardDensity(case['densities'])
    assert result == pytest.approx(expected, rel = 2e-6 if precision == 'float32' else 1e-12)


@pytest.mark.parametrize('precision', precisions)
@pytest.mark.parametrize('dimensions, periodic', cases)
@pytest.mark.parametrize('kernelType', kernelTypes)
def testXsphMatchesTheOracle(precision, dimensions, periodic, kernelType):
    # tests/testSphGpu.py:362-375: the fluid to 1e-5 of the largest value, the boundary unchanged
    print(f'xsph {precision} {dimensions}D periodic={periodic} {kernelType}', flush = True)
    case = cloud(dimensions, periodic, kernelType = kernelType)
    fluid = case['isFluid']
    expected = oracle.xsphVelocity(physicsFor(case), case['positions'], case['velocities'], case['densities'],
                                   case['masses'], fluid, 0.5)
    result = pairSums(case, precision).xsphVelocity(case['velocities'], case['densities'], 0.5)
    assert result[fluid] == pytest.approx(expected[fluid], abs = tolerance[precision] * np.abs(expected).max())
    rounded = case['velocities'][~fluid].astype(np.float32) if precision == 'float32' else case['velocities'][~fluid]
    assert np.array_equal(result[~fluid], rounded)


@pytest.mark.parametrize('precision', precisions)
def testShepardAndXsphWithMixedLengthsMatchTheOracle(precision):
    # tests/testSphRefinement.py:307-326: h_ij = (h_i + h_j) / 2, each particle's own W(0, h_i)
    case = cloud(3, False)
    rng = np.random.default_rng(5)
    lengths = case['h'] * np.where(rng.uniform(size = len(case['positions'])) < 0.3, 0.5, 1.0)
    physics = physicsFor(case, hmin = float(lengths.min()))
    fluid = case['isFluid']
    sums = pairSums(case, precision, lengths)
    shepard = oracle.shepardDensity(physics, case['positions'], case['densities'], case['masses'], fluid, lengths)
    assert sums.shepardDensity(case['densities']) == pytest.approx(shepard,
                                                                   rel = 2e-6 if precision == 'float32' else 1e-12)
    xsph = oracle.xsphVelocity(physics, case['positions'], case['velocities'], case['densities'], case['masses'], fluid,
                               0.5, lengths)
    result = sums.xsphVelocity(case['velocities'], case['densities'], 0.5)
    assert result[fluid] == pytest.approx(xsph[fluid], abs = tolerance[precision] * np.abs(xsph).max())


def testXsphRefusesAnEpsilonOutsideZeroToOne():
    case = cloud(2, False)
    sums = pairSums(case, 'float32')
    with pytest.raises(ValueError, match = 'epsilon'):
        sums.xsphVelocity(case['velocities'], case['densities'], 1.5)


@pytest.mark.parametrize('precision', precisions)
def testDiffusionDifferencesKeepTheirDigits(precision):
    # tests/testSphGpu.py:284-297: a density gradient of 1e-4 kg/m^3 a spacing survives float32
    case = cloud(2, False)
    densities = rho0 + 1e-4 * case['positions'][:, 0] / 0.04
    physics = physicsFor(case, diffusion = 'molteni', delta = 0.1)
    still = np.zeros_like(case['velocities'])
    expected = oracle.continuityRate(physics, case['positions'], still, densities, case['masses'], case['isFluid'])
    sums = pairSums(case, precision)
    sums.setDiffusion('molteni', 0.1, c0, rho0, 9.81)
    result = sums.continuityRate(still, densities)
    assert result == pytest.approx(expected, abs = 1e-3 * np.abs(expected).max())


@pytest.mark.parametrize('precision', precisions)
@pytest.mark.parametrize('dimensions, periodic', cases)
def testNeighborSetsAreBitExact(precision, dimensions, periodic):
    # Each precision's predicate, evaluated by the device and by the host restatement, selects the
    # same neighbors; the float64 set is also the brute-force set
    print(f'neighbors {precision} {dimensions}D periodic={periodic}', flush = True)
    case = cloud(dimensions, periodic)
    lengths = np.full(len(case['positions']), case['h'])
    offsets, rows = pairSums(case, precision).neighborLists()
    expectedOffsets, expectedRows = oracle.neighbors(physicsFor(case), case['positions'], lengths, precision)
    assert np.array_equal(offsets, expectedOffsets)
    assert np.array_equal(rows, expectedRows)


@pytest.mark.parametrize('precision', precisions)
@pytest.mark.parametrize('dimensions', [2, 3])
@pytest.mark.parametrize('placement', ['straddling the origin', 'partly beyond the grid', 'beyond the grid',
                                       'kilometres out'])
def testRowsCutAlongXKeepEveryNeighbor(precision, dimensions, placement):
    # 3D rows are cut along x by the particle's distance across to each row, from its unwrapped
    # cell, in whole eighths of each cell and then by a search within the first and last eighth, and
    # a cell holding a particle from outside the grid along x is taken whole; 2D rows are walked
    # whole: a cloud across the origin (cell coordinates of both signs) and ones given a grid
    # narrower than themselves (cells holding particles of several wrapped cells, some or most of
    # them) keep the brute-force neighbor sets. Kilometres from the origin a float32 point's
    # remainder is a large share of a spacing, the case where the 3D test on points alone must hand
    # the most pairs to the exact test
    print(f'rows cut along x {dimensions}D {placement} {precision}', flush = True)
    case = cloud(dimensions)
    positions = case['positions'] - 25.0 - 0.5 * (case['positions'].max(axis = 0) - 25.0)
    if placement == 'kilometres out':
        positions = positions + np.array([4000.0, -2000.0, 1000.0][:dimensions])
    from pysurf_amd.pair_sums import PairSums
    low, high = positions.min(axis = 0), positions.max(axis = 0)
    if placement in ('partly beyond the grid', 'beyond the grid'):
        high = low + (0.4 if placement == 'beyond the grid' else 0.8) * (high - low)
    sums = PairSums(dimensions, case['h'], low, high, 0.0, 'float64', 'cuda:0', case['kernelType'],
                    pairPrecision = precision)
    sums.update(positions, case['masses'], case['isFluid'], None)
    offsets, rows = sums.neighborLists()
    expectedOffsets, expectedRows = oracle.neighbors(physicsFor(dict(case, positions = positions)), positions,
                                                     np.full(len(positions), case['h']), precision)
    assert np.array_equal(offsets, expectedOffsets)
    assert np.array_equal(rows, expectedRows)
    # the float64 sets are the brute-force sets
    if precision == 'float64' and placement != 'kilometres out':
        assert len(rows) == 2 * len(case['i'])


@pytest.mark.parametrize('precision', precisions)
def testMixedLengthsMatchTheOracle(precision):
    # Particles of two smoothing lengths pair at h_ij = (h_i + h_j) / 2 (sphGpu.py:52-56)
    case = cloud(2, False, kernelType = 'wendlandC2')
    rng = np.random.default_rng(5)
    lengths = case['h'] * np.where(rng.uniform(size = len(case['positions'])) < 0.3, 0.5, 1.0)
    physics = physicsFor(case, boundaryFactor = 0.3, hmin = float(lengths.min()))
    expected = oracle.accelerations(physics, case['positions'], case['velocities'], case['densities'],
                                    case['pressures'], case['masses'], case['isFluid'], lengths = lengths)
    sums = pairSums(case, precision, lengths)
    result = sums.pairAccelerations(case['velocities'], case['densities'], case['pressures'], c0, alpha, 0.3,
                                    gravityOf(2))
    assert result == pytest.approx(expected, abs = tolerance[precision] * np.abs(expected).max())
    offsets, rows = sums.neighborLists()
    expectedOffsets, expectedRows = oracle.neighbors(physics, case['positions'], lengths, precision)
    assert np.array_equal(rows, expectedRows) and np.array_equal(offsets, expectedOffsets)


def testTheEngineFailsLoudlyWithoutADevicePath():
    from pysurf_amd.pair_sums import PairSums
    with pytest.raises(ValueError):
        PairSums(2, 0.05, np.zeros(2), np.ones(2), device = 'cpu')


def testHostStepSumsRefuseRunsWithOtherOptions():
    # The host-step sums launch by the first run's options, so a context whose runs differ in the
    # limiter, the correction or the diffusion is refused rather than summed with run 0's options
    import dataclasses

    from pysurf_amd import _native
    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.scenarios import solitaryOnSlope
    first = solitaryOnSlope(particlesPerDepth = 6, seed = 1)
    for change in ({'limiter': not first.limiter}, {'correction': not first.correction}, {'diffusion': 'molteni'}):
        batch = BatchEngine([first, dataclasses.replace(solitaryOnSlope(particlesPerDepth = 6, seed = 2), **change)],
                            'float32')
        engine = batch.engine
        status = engine.library.sph_velocity_gradients(engine.handle, engine.pointer('velocity'),
                                                       engine.pointer('density'), engine.pointer('matrix'))
        assert status == _native.SPH_ERR_VALUE, change
        assert b'share' in engine.library.sph_last_error(engine.handle)
    # runs that agree are summed, over the cells of the bound positions
    batch = BatchEngine([first, solitaryOnSlope(particlesPerDepth = 6, seed = 2)], 'float32')
    engine = batch.engine
    assert engine.library.sph_build_cells(engine.handle) == _native.SPH_OK
    assert engine.library.sph_velocity_gradients(engine.handle, engine.pointer('velocity'), engine.pointer('density'),
                                                 engine.pointer('matrix')) == _native.SPH_OK


def testWhichMatricesAStepNeeds():
    # needsCorrections: the continuity correction or the 'antuono' diffusion; needsVelocityGradients:
    # the viscosity limiter or the closure (sphGpu.py:1659-1667); the host steps refuse G and L when
    # neither is on
    from pysurf_amd.pair_sums import PairSums
    sums = PairSums(2, 0.05, np.zeros(2), np.ones(2))
    assert not sums.needsCorrections and not sums.needsVelocityGradients
    sums.setContinuityCorrection(True)
    assert sums.needsCorrections
    sums.setContinuityCorrection(False)
    sums.setDiffusion('antuono', 0.1, 20.0, 1000.0, 9.81)
    assert sums.needsCorrections
    sums.setViscosityLimiter(True)
    assert sums.needsVelocityGradients
    sums.setViscosityLimiter(False)
    sums.setClosure(True, soundSpeed = 20.0)
    assert sums.needsVelocityGradients
    sums.setClosure(False)
    assert not sums.needsVelocityGradients
    with pytest.raises(ValueError):
        sums.velocityGradients(np.zeros((0, 2)), np.zeros(0))
