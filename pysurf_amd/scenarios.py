'''

The runs the engine is measured and checked on, built as pySurf builds its 2D breaking waves
(BreakingWave.build, src/pysurf/BreakingWave.py:658-767): a plane beach, a bed boundary with an
offshore wall, a first-order solitary wave laid on the lattice, jittered, its densities
hydrostatic beneath the wave (WcsphSolver.initialize, sphSolver.py:1869-1897), and the
delta-SPH scheme's options (BreakingWave.py:722-725).

'''

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from pysurf_amd.bathymetry import Bathymetry, HeightField
from pysurf_amd.boundaries import BedBoundary, ContainerBoundary, HeightFieldBoundary
from pysurf_amd.particles import ParticleSystem, jitterPoints
from pysurf_amd.refinement import splitParticles
from pysurf_amd.waves import freshwaterDensity, gravity, solitaryWaveFluid, solitaryWaveFluid3D, solitaryWaveProfile


def taitDensity(pressure, referenceDensity: float, soundSpeed: float):
    '''rho0 (1 + p / B)^(1/7), B = rho0 c0^2 / 7 (sphSolver.py:559-566).'''
    stiffness = referenceDensity * soundSpeed**2 / 7.0
    return referenceDensity * (1.0 + np.asarray(pressure) / stiffness)**(1.0 / 7.0)


@dataclass
class Case:

    '''One run ready for the engine: particles, boundary, grid and the scheme's physics.'''

    name: str
    particles: ParticleSystem
    boundary: BedBoundary
    dimensions: int
    spacing: float
    smoothingLength: float
    domainMin: np.ndarray
    domainMax: np.ndarray
    kernelType: str = 'wendlandC2'
    referenceDensity: float = freshwaterDensity
    soundSpeed: float = 10.0 * math.sqrt(gravity)
    alpha: float = 0.01
    boundaryViscosityFactor: float = 0.0
    gravity: np.ndarray = field(default_factory = lambda: np.array([0.0, -gravity]))
    diffusion: str = 'hydrostatic'
    delta: float = 0.1
    extrapolatedWalls: bool = True
    negativePressure: bool = False
    correction: bool = True
    limiter: bool = True
    densityFloor: float = 0.9 * freshwaterDensity
    cfl: float = 0.25
    maxTimeStep: float = 1.0
    filterInterval: int = 0      # the Shepard filter every this many steps (densityFilterInterval); 0 never
    parameters: dict = field(default_factory = dict)

    @property
    def fluidCount(self) -> int:
        return self.particles.nFluid

    @property
    def count(self) -> int:
        return self.particles.nParticles


def solitaryOnSlope(particlesPerDepth: int = 16, depth: float = 1.0, waveHeight: float = 0.45, crestX: float = 4.0,
                    slope: float = 1.0 / 15.0, toeX: float = 8.0, length: float = 25.0, jitter: float = 1e-3,
                    seed: int = 0, smoothingRatio: float = 1.3, kernelType: str = 'wendlandC2',
                    soundSpeedFactor: float = 10.0, name: str | None = None) -> Case:

    '''

    A first-order sech^2 solitary wave of H = `waveHeight` at x = `crestX` over a bed flat to
    `toeX`, then a plane slope, `length` long; the sound speed `soundSpeedFactor` sqrt(g d); each
    fluid point jittered by up to `jitter` spacings with `seed`.

    '''

    d, s = depth, depth / particlesPerDepth
    layers = max(3, math.ceil(2.0 * smoothingRatio - 1e-9))          # BreakingWave.py:584-586
    wallTop = d + 3.0 * waveHeight                                   # BreakingWave.py:667
    bed = Bathymetry.planeBeach(depth = d, toeX = toeX, slope = slope, endX = length)
    endWallTop = wallTop if bed.elevation(bed.endX) < d else None
    boundary = BedBoundary(bed, s, nLayers = layers, wallTop = wallTop, endWallTop = endWallTop)
    fluid, velocities = solitaryWaveFluid(bed, waveHeight, crestX, s)
    if jitter > 0.0:
        surface = lambda x: d + solitaryWaveProfile(x, waveHeight, d, crestX)
        fluid = jitterPoints(fluid, s, jitter, seed, [bed.startX], [bed.endX], bed.elevation, surface)
    particles = ParticleSystem.assemble(fluid, boundary.positions(), s, freshwaterDensity, velocities)
    c0 = soundSpeedFactor * math.sqrt(gravity * d)
    x = particles.positions[:, 0]
    surfaceElevation = d + solitaryWaveProfile(x, waveHeight, d, crestX)
    depthBelow = np.maximum(surfaceElevation - particles.positions[:, -1], 0.0)
    particles.densities = taitDensity(freshwaterDensity * gravity * depthBelow, freshwaterDensity, c0)
    h = smoothingRatio * s
    return Case(name = name or f'solitary-{particlesPerDepth}ppd', particles = particles, boundary = boundary,
                dimensions = 2, spacing = s, smoothingLength = h,
                domainMin = np.array([-(layers + 1.0) * s, -(layers + 1.0) * s]),
                domainMax = np.array([bed.endX + (layers + 1.0) * s, wallTop]), kernelType = kernelType,
                soundSpeed = c0, densityFloor = 0.9 * freshwaterDensity,
                parameters = dict(particlesPerDepth = particlesPerDepth, waveHeight = waveHeight, slope = slope,
                                  crestX = crestX, toeX = toeX, length = length, seed = seed, depth = d))


def stillTank(spacing: float = 0.005, width: float = 0.2, height: float = 0.15, depth: float = 0.075,
              jitter: float = 0.0, seed: int = 0, smoothingRatio: float = 1.3, kernelType: str = 'cubicSpline',
              **physics) -> Case:

    '''

    Water at rest, `depth` deep, in an open box `width` x `height` (ContainerBoundary, 3 layers) at
    hydrostatic densities, c0 = 10 sqrt(g height); each fluid point jittered by up to `jitter`
    spacings with `seed`. By default the reference's small tank of its at-rest checks
    (tests/testSphGpu.py:440-457, 557-575: 0.2 x 0.15 m half full, 5 mm spacing, cubic spline),
    with `physics` setting the scheme (Case's fields).

    '''

    s = spacing
    layers = 3
    boundary = ContainerBoundary([0.0, 0.0], [width, height], s, nLayers = layers, openTop = True)
    xs = np.arange(0.5 * s, width, s)
    zs = np.arange(0.5 * s, depth, s)
    fluid = np.array([(x, z) for z in zs for x in xs])
    if jitter > 0.0:
        fluid = jitterPoints(fluid, s, jitter, seed, [0.0], [width], lambda x: np.zeros_like(x),
                             lambda x: np.full_like(x, depth))
    particles = ParticleSystem.assemble(fluid, boundary.positions(), s, freshwaterDensity, np.zeros_like(fluid))
    c0 = 10.0 * math.sqrt(gravity * height)
    depthBelow = np.maximum(depth - particles.positions[:, -1], 0.0)
    particles.densities = taitDensity(freshwaterDensity * gravity * depthBelow, freshwaterDensity, c0)
    options = dict(kernelType = kernelType, soundSpeed = c0, densityFloor = 0.9 * freshwaterDensity)
    options.update(physics)
    return Case(name = 'tank', particles = particles, boundary = boundary, dimensions = 2, spacing = s,
                smoothingLength = smoothingRatio * s, domainMin = np.array([-(layers + 1.0) * s, -(layers + 1.0) * s]),
                domainMax = np.array([width + (layers + 1.0) * s, height]),
                parameters = dict(depth = depth, width = width, height = height, seed = seed), **options)


def singleRun(seed: int = 0, particlesPerDepth: int = 64) -> Case:
    '''The 2D solitary wave at 64 particles per depth with H/d ~ U[0.3, 0.6] and slope ~ U[1/35, 1/10] from `seed`.'''
    rng = np.random.default_rng(seed)
    waveHeight = float(rng.uniform(0.3, 0.6))
    slope = float(rng.uniform(1.0 / 35.0, 1.0 / 10.0))
    return solitaryOnSlope(particlesPerDepth = particlesPerDepth, waveHeight = waveHeight, slope = slope, seed = seed,
                           name = f'solitary-{particlesPerDepth}ppd-seed{seed}')


def sweepRun(seed: int, resolutions = (16, 24, 32, 48)) -> Case:
    '''One run of the batched sweep: slope U[1/35, 1/8], H/d U[0.2, 0.6], ppd from `resolutions`, length U[20, 30] d.'''
    rng = np.random.default_rng(1000 + seed)
    slope = float(rng.uniform(1.0 / 35.0, 1.0 / 8.0))
    waveHeight = float(rng.uniform(0.2, 0.6))
    ppd = int(rng.choice(resolutions))
    length = float(rng.uniform(20.0, 30.0))
    toeX = min(8.0, 0.5 * length)
    return solitaryOnSlope(particlesPerDepth = ppd, waveHeight = waveHeight, slope = slope, length = length,
                           toeX = toeX, seed = seed, name = f'sweep-{seed}')


def exactSolitaryOnSlope(particlesPerDepth: int = 64, depth: float = 1.0, waveHeight: float = 0.3,
                         slope: float = 1.0 / 35.0, runupAllowance: float = 0.8, smoothingRatio: float = 3.0,
                         bedFitted: bool = True, jitter: float = 0.0, seed: int = 1, alpha: float = 0.01,
                         soundSpeedFactor: float = 10.0, faceArc: float = 2.0, name: str | None = None) -> Case:

    '''

    The δ-SPH start of pySurf's BreakingWave (BreakingWave.py:651-765, scheme 'deltaSph', starting
    wave 'exact'): the exact solitary wave of H = `waveHeight` laid in bed-fitted columns (or on
    the lattice) a half length from the offshore wall, a plane beach from a half length past the
    crest to (1 + `runupAllowance`) depths of rise past the still shoreline, the Wendland kernel at
    h = `smoothingRatio` spacings with six wall layers, every particle's density from the wave's
    pressure (each wall particle's from the water just inside it), and the smoothed-surface tracker
    from the toe. The defaults are Grilli's 1:35 case (validation/gpuSph/grilliBreaking.py).

    '''

    from pysurf_amd.exact_wave import exactSolitaryBoundaryPressure, exactSolitaryWaveFluid
    from pysurf_amd.waves import solitaryWaveHalfLength, solitaryWaveSpeed
    d, s, rho = depth, depth / particlesPerDepth, freshwaterDensity
    halfLength = solitaryWaveHalfLength(waveHeight, d, 0.01)
    crestX = halfLength
    toeX = crestX + halfLength
    endX = toeX + (1.0 + runupAllowance) * d / slope
    layers = max(3, math.ceil(2.0 * smoothingRatio - 1e-9))
    wallTop = d + 3.0 * waveHeight
    bed = Bathymetry.planeBeach(depth = d, toeX = toeX, slope = slope, endX = endX)
    boundary = BedBoundary(bed, s, nLayers = layers, wallTop = wallTop, endWallTop = None)
    fluid, velocities, pressures = exactSolitaryWaveFluid(bed, waveHeight, crestX, s, bedFitted = bedFitted,
                                                          jitter = jitter, seed = seed)
    particles = ParticleSystem.assemble(fluid, boundary.positions(), s, rho, velocities)
    c0 = soundSpeedFactor * solitaryWaveSpeed(waveHeight, d)
    walls = exactSolitaryBoundaryPressure(bed, particles.positions[~particles.isFluid], waveHeight, crestX, s, gravity)
    particles.densities[particles.isFluid] = taitDensity(rho * pressures, rho, c0)
    particles.densities[~particles.isFluid] = taitDensity(rho * walls, rho, c0)
    return Case(name = name or f'exact-solitary-{particlesPerDepth}ppd', particles = particles, boundary = boundary,
                dimensions = 2, spacing = s, smoothingLength = smoothingRatio * s,
                domainMin = np.array([-(layers + 1.0) * s, -(layers + 1.0) * s]),
                domainMax = np.array([endX + (layers + 1.0) * s, wallTop]), kernelType = 'wendlandC2',
                soundSpeed = c0, alpha = alpha, densityFloor = 0.9 * rho,
                parameters = dict(particlesPerDepth = particlesPerDepth, waveHeight = waveHeight, slope = slope,
                                  crestX = crestX, toeX = toeX, length = endX, seed = seed, depth = d,
                                  faceMeasure = 'contour', faceArc = faceArc, start = 'exact'))


def ridge3D(particlesPerDepth: int = 32, depth: float = 1.0, waveHeight: float = 0.5, crestX: float = 4.0,
            ridgeHeight: float = 0.5, ridgeWidth: float = 2.0, slope: float = 1.0 / 15.0, toeX: float = 6.0,
            length: float = 20.0, width: float = 10.0, obliquity: float = 25.0, zoneHalfWidth: float | None = 0.75,
            lengthRatio: float = 1.0, jitter: float = 1e-3, seed: int = 0, smoothingRatio: float = 1.3,
            kernelType: str = 'wendlandC2', soundSpeedFactor: float = 10.0, start: str = 'first-order',
            name: str | None = None) -> Case:

    '''

    A 3D solitary wave of H = `waveHeight` with its crest line along y at x = `crestX`, over a
    Gaussian ridge (HeightField.gaussianRidge: `ridgeHeight` tall, `ridgeWidth` wide, rising
    shoreward from `toeX` at `slope`, its axis turned `obliquity` degrees from the wave's direction)
    in a walled tank `length` x `width` depths. Fluid within `zoneHalfWidth` depths of the crest is
    split once into 8 daughters (keeping h with `lengthRatio` 1); None leaves it whole. Lengths are
    in units of `depth`. `start` 'first-order' lays the sech^2 wave with hydrostatic densities;
    'exact' lays the exact wave in bed-fitted columns (jittered by `jitter` spacings), every
    particle's density from its pressure and each wall particle's from the water just inside it
    (BreakingWave3D.py:660-690).

    '''

    d, s = depth, depth / particlesPerDepth
    layers = max(3, math.ceil(2.0 * smoothingRatio - 1e-9))
    wallTop = d + 2.0 * waveHeight
    field = HeightField.gaussianRidge(depth = d, height = ridgeHeight, width = ridgeWidth, slope = slope, toeX = toeX,
                                      length = length, halfWidth = 0.5 * width, direction = obliquity)
    boundary = HeightFieldBoundary(field, s, nLayers = layers, wallTop = wallTop)
    startX, endX, startY, endY = field.bounds
    if start == 'exact':
        return exactRidge(field, boundary, d, s, waveHeight, crestX, layers, wallTop, zoneHalfWidth, lengthRatio, jitter,
                          seed, smoothingRatio, kernelType, soundSpeedFactor, name,
                          dict(particlesPerDepth = particlesPerDepth, waveHeight = waveHeight, crestX = crestX,
                               ridgeHeight = ridgeHeight, ridgeWidth = ridgeWidth, slope = slope, toeX = toeX,
                               length = length, width = width, obliquity = obliquity, zoneHalfWidth = zoneHalfWidth,
                               lengthRatio = lengthRatio, seed = seed, depth = d, start = 'exact'))
    if start != 'first-order':
        raise ValueError(f"The start is 'first-order' or 'exact', not {start!r}.")
    fluid, velocities = solitaryWaveFluid3D(field, waveHeight, crestX * d, s)
    if jitter > 0.0:
        surface = lambda x, y: d + solitaryWaveProfile(x, waveHeight, d, crestX * d)
        fluid = jitterPoints(fluid, s, jitter, seed, [startX, startY], [endX, endY], field.elevation, surface)
    particles = ParticleSystem.assemble(fluid, boundary.positions(), s, freshwaterDensity, velocities)
    h = smoothingRatio * s
    if zoneHalfWidth is not None:
        zone = np.abs(particles.positions[:, 0] - crestX * d) < zoneHalfWidth * d
        splitParticles(particles, zone, freshwaterDensity, h, lengthRatio)
    c0 = soundSpeedFactor * math.sqrt(gravity * d)
    surfaceElevation = d + solitaryWaveProfile(particles.positions[:, 0], waveHeight, d, crestX * d)
    depthBelow = np.maximum(surfaceElevation - particles.positions[:, -1], 0.0)
    particles.densities = taitDensity(freshwaterDensity * gravity * depthBelow, freshwaterDensity, c0)
    margin = (layers + 1.0) * s
    return Case(name = name or f'ridge3d-{particlesPerDepth}ppd', particles = particles, boundary = boundary,
                dimensions = 3, spacing = s, smoothingLength = h,
                domainMin = np.array([startX - margin, startY - margin, field.lowestElevation() - (layers + 2.0) * s]),
                domainMax = np.array([endX + margin, endY + margin, wallTop]), kernelType = kernelType,
                soundSpeed = c0, gravity = np.array([0.0, 0.0, -gravity]), densityFloor = 0.9 * freshwaterDensity,
                parameters = dict(particlesPerDepth = particlesPerDepth, waveHeight = waveHeight, crestX = crestX,
                                  ridgeHeight = ridgeHeight, ridgeWidth = ridgeWidth, slope = slope, toeX = toeX,
                                  length = length, width = width, obliquity = obliquity,
                                  zoneHalfWidth = zoneHalfWidth, lengthRatio = lengthRatio, seed = seed, depth = d))


def exactRidge(field, boundary, d: float, s: float, waveHeight: float, crestX: float, layers: int, wallTop: float,
               zoneHalfWidth, lengthRatio: float, jitter: float, seed: int, smoothingRatio: float, kernelType: str,
               soundSpeedFactor: float, name, parameters: dict) -> Case:
    '''ridge3D's exact start: the exact wave's flow and pressure in bed-fitted columns, the walls' pressures from the water inside them.'''
    from pysurf_amd.exact_wave import exactSolitaryBoundaryPressure3D, exactSolitaryWaveFluid3D
    from pysurf_amd.waves import solitaryWaveSpeed
    rho = freshwaterDensity
    fluid, velocities, pressures = exactSolitaryWaveFluid3D(field, waveHeight, d, crestX * d, s, bedFitted = True,
                                                            jitter = jitter, seed = seed)
    particles = ParticleSystem.assemble(fluid, boundary.positions(), s, rho, velocities)
    fluidPressure = np.zeros(particles.nParticles)
    fluidPressure[particles.isFluid] = pressures
    h = smoothingRatio * s
    if zoneHalfWidth is not None:
        zone = np.abs(particles.positions[:, 0] - crestX * d) < zoneHalfWidth * d
        count = particles.nParticles
        splitParticles(particles, zone, rho, h, lengthRatio)
        # A daughter carries its mother's pressure: the first takes the mother's slot, the others
        # follow mother by mother
        mothers = np.flatnonzero(zone & particles.isFluid[:count])
        daughters = (particles.nParticles - count) // max(len(mothers), 1)
        fluidPressure = np.concatenate([fluidPressure, np.repeat(fluidPressure[mothers], daughters)])
    c0 = soundSpeedFactor * solitaryWaveSpeed(waveHeight, d)
    walls = exactSolitaryBoundaryPressure3D(field, particles.positions[~particles.isFluid], waveHeight, d, crestX * d, s,
                                            gravity)
    particles.densities[particles.isFluid] = taitDensity(rho * fluidPressure[particles.isFluid], rho, c0)
    particles.densities[~particles.isFluid] = taitDensity(rho * walls, rho, c0)
    startX, endX, startY, endY = field.bounds
    margin = (layers + 1.0) * s
    return Case(name = name or f'exact-ridge3d-{parameters["particlesPerDepth"]}ppd', particles = particles,
                boundary = boundary, dimensions = 3, spacing = s, smoothingLength = h,
                domainMin = np.array([startX - margin, startY - margin, field.lowestElevation() - (layers + 2.0) * s]),
                domainMax = np.array([endX + margin, endY + margin, wallTop]), kernelType = kernelType,
                soundSpeed = c0, gravity = np.array([0.0, 0.0, -gravity]), densityFloor = 0.9 * rho,
                parameters = parameters)


def servingRidge(seed: int) -> Case:
    '''A small 3D ridge run of the serving mix at 16 ppd: tank 16-24 d long, 6-12 d wide; ridge 0.3-0.6 d tall, 1.5-3 d wide, turned 0-30 degrees.'''
    rng = np.random.default_rng(5000 + seed)
    length = float(rng.uniform(16.0, 24.0))
    return ridge3D(particlesPerDepth = 16, waveHeight = float(rng.uniform(0.3, 0.5)), crestX = 4.0,
                   ridgeHeight = float(rng.uniform(0.3, 0.6)), ridgeWidth = float(rng.uniform(1.5, 3.0)),
                   toeX = float(rng.uniform(6.0, 8.0)), length = length, width = float(rng.uniform(6.0, 12.0)),
                   obliquity = float(rng.uniform(0.0, 30.0)), zoneHalfWidth = None, seed = seed,
                   name = f'serving-ridge-{seed}')


def servingMix(count: int, seed: int = 0, share3D: float = 0.1) -> list:
    '''`count` runs of the serving stream: a `share3D` share of small 3D ridge runs, the rest 2D runs of the batched sweep.'''
    rng = np.random.default_rng(7000 + seed)
    kinds = rng.uniform(size = count) < share3D
    return [servingRidge(1000 * seed + k) if three else sweepRun(1000 * seed + k) for k, three in enumerate(kinds)]
