'''

The solver shell over the HIP engine: pySurf's `SimulationConfig` and `WcsphSolver`
(src/pysurf/sphSolver.py:288-484, 1430-2248) with the backend the engine provides, so a caller
that builds a config and a solver, as `BreakingWave.build` does (BreakingWave.py:722-767), runs
its steps on the MI355X.

`backend` takes 'hip' (and 'gpu', the reference's name for its device backend, as an alias). Every
step runs on the device: `advance` keeps the state and the time step there and replays the steps
as a HIP graph; `step` is one device step with the time step taken on the host, as the reference's
`advance` loop takes it (sphSolver.py:1980-1986).

'''

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from pysurf_amd.errors import NumericalInstabilityError
from pysurf_amd.pair_sums import PairSums
from pysurf_amd.particles import ParticleSystem
from pysurf_amd.waves import gravity as standardGravity

taitExponent = 7.0


def elasticEnergy(densities, referenceDensity: float, soundSpeed: float, tension: bool = False) -> np.ndarray:
    '''

    The elastic energy per unit mass the Tait fluid stores at each density [J/kg]: the integral of
    p / s^2 from rho0, (c0^2 / 7) x^2 (21 + 35 x + 35 x^2 + 21 x^3 + 7 x^4 + x^5) / (6 (1 + x)) with
    x = rho / rho0 - 1, zero below rho0 unless the fluid carries `tension` (sphEnergy.py:116-160).

    '''
    x = np.asarray(densities, dtype = float) / referenceDensity - 1.0
    polynomial = 21.0 + x * (35.0 + x * (35.0 + x * (21.0 + x * (7.0 + x))))
    energy = soundSpeed**2 / 7.0 * x * x * polynomial / (6.0 * (1.0 + x))
    return energy if tension else np.where(x < 0.0, 0.0, energy)


@dataclass
class SimulationConfig:

    '''A simulation's configuration, field for field the reference's (sphSolver.py:288-440).'''

    domainMin: np.ndarray
    domainMax: np.ndarray
    particleSpacing: float = 0.005
    smoothingLengthRatio: float = 1.3
    referenceDensity: float = 1000.0
    gravity: np.ndarray = field(default_factory = lambda: np.array([0.0, -standardGravity]))
    endTime: float = 3.0
    maxTimeStep: float = 1e-4
    outputInterval: float = 0.02
    dimensions: int = 2
    kernelType: str = 'cubicSpline'
    speedOfSound: float | None = None
    speedOfSoundFactor: float = 10.0
    alphaViscosity: float = 0.1
    boundaryViscosityFactor: float = 0.0
    xsphEpsilon: float = 0.5
    cflNumber: float = 0.25
    densityFilterInterval: int = 30
    densityFloor: float = 0.0
    initialSurface: object = None
    periodicWidth: float | None = None
    densityDiffusion: str = 'none'
    diffusionDelta: float = 0.1
    particleShifting: bool = False
    shiftingSpeed: str = 'reference'
    shiftingMach: float = 0.1
    wallPressure: str = 'continuity'
    negativePressure: bool = False
    continuityCorrection: bool = False
    viscosityLimiter: bool = False
    turbulenceClosure: str = 'none'
    backend: str = 'hip'
    pairPrecision: str = 'float32'

    @property
    def smoothingLength(self) -> float:
        return self.smoothingLengthRatio * self.particleSpacing

    @property
    def supportRadius(self) -> float:
        return 2.0 * self.smoothingLength

    @property
    def domainSize(self) -> np.ndarray:
        return np.asarray(self.domainMax) - np.asarray(self.domainMin)


@dataclass
class SimulationState:

    '''Scalar diagnostics of one instant (sphSolver.py:486-521).'''

    time: float
    step: int
    dt: float
    kineticEnergy: float
    potentialEnergy: float
    maxVelocity: float
    maxDensityError: float

    @property
    def totalEnergy(self) -> float:
        return self.kineticEnergy + self.potentialEnergy


def taitEquationOfState(density, referenceDensity: float, soundSpeed: float, clampNegative: bool = True):
    '''p = B ((rho / rho0)^7 - 1), clamped at zero by default (sphSolver.py:541-557).'''
    stiffness = referenceDensity * soundSpeed**2 / taitExponent
    pressure = stiffness * ((np.asarray(density) / referenceDensity)**taitExponent - 1.0)
    return np.maximum(pressure, 0.0) if clampNegative else pressure


def taitDensity(pressure, referenceDensity: float, soundSpeed: float):
    '''The inverse of the Tait equation (sphSolver.py:559-566).'''
    stiffness = referenceDensity * soundSpeed**2 / taitExponent
    return referenceDensity * (1.0 + np.asarray(pressure) / stiffness)**(1.0 / taitExponent)


def hydrostaticDensity(depthBelowSurface, referenceDensity: float, soundSpeed: float, gravity: float = standardGravity):
    '''The density carrying rho0 g z at depth z (sphSolver.py:1108-1120).'''
    pressure = referenceDensity * gravity * np.maximum(np.asarray(depthBelowSurface, dtype = float), 0.0)
    return taitDensity(pressure, referenceDensity, soundSpeed)


def speedOfSound(characteristicDepth: float, factor: float = 10.0, gravity: float = standardGravity) -> float:
    '''c0 = factor sqrt(g H) (sphSolver.py:527-539).'''
    return factor * math.sqrt(gravity * characteristicDepth)


def cellOrderKey(positions, cellWidth: float, shape) -> np.ndarray:
    '''Cells from the origin wrapped onto the grid, x fastest (sphSolver.py:1379-1408).'''
    cells = np.floor(np.asarray(positions, dtype = float) / cellWidth).astype(np.int64)
    nx, ny, nz = (int(count) for count in shape)
    x, y = np.mod(cells[:, 0], nx), np.mod(cells[:, 1], ny)
    z = np.mod(cells[:, 2], nz) if cells.shape[1] == 3 else np.zeros_like(x)
    return (z * ny + y) * nx + x


def cflTimeStep(smoothingLength: float, soundSpeed: float, maxSpeed: float, maxAcceleration: float,
                cflNumber: float, maxTimeStep: float) -> float:
    '''min(CFL h / (c0 + |v|max), CFL sqrt(h / |a|max), maxTimeStep) (sphSolver.py:1410-1424).'''
    timeStep = cflNumber * smoothingLength / (soundSpeed + maxSpeed)
    if maxAcceleration > 1e-12:
        timeStep = min(timeStep, cflNumber * math.sqrt(smoothingLength / maxAcceleration))
    return min(timeStep, maxTimeStep)


class WcsphSolver:

    '''

    The WCSPH time-stepping loop on the HIP engine (sphSolver.py:1430-2248).

    Parameters:
    -----------
    config : SimulationConfig
        Domain, resolution and numerical parameters
    boundary : ContainerBoundary | BedBoundary | HeightFieldBoundary | None
        The boundary the device enforces after each step

    '''

    def __init__(self, config: SimulationConfig, kernel = None, boundary = None) -> None:
        if config.periodicWidth is not None and config.periodicWidth <= 2.0 * config.supportRadius:
            raise ValueError(f'The periodic width {config.periodicWidth} m must exceed twice the kernel support, '
                             f'{2.0 * config.supportRadius} m.')
        if config.backend not in ('hip', 'gpu'):
            raise ValueError(f"Backend '{config.backend}' is not 'hip' (or its alias 'gpu').")
        if config.densityDiffusion not in ('none', 'molteni', 'hydrostatic', 'antuono'):
            raise ValueError(f"Density diffusion '{config.densityDiffusion}' is not 'none', 'molteni', 'hydrostatic' "
                             "or 'antuono'.")
        if config.wallPressure not in ('continuity', 'extrapolated'):
            raise ValueError(f"Wall pressure '{config.wallPressure}' is not 'continuity' or 'extrapolated'.")
        if config.turbulenceClosure not in ('none', 'deltaLes'):
            raise ValueError(f"Turbulence closure '{config.turbulenceClosure}' is not 'none' or 'deltaLes'.")
        if config.densityFilterInterval < 0:
            raise ValueError('The density filter interval cannot be negative.')
        if config.xsphEpsilon > 0.0:
            raise ValueError('The device step does not move the fluid with XSPH (PairSums.xsphVelocity gives it as a '
                             'host step); the delta-SPH scheme sets xsphEpsilon to 0.')
        if getattr(config, 'particleShifting', False):
            raise ValueError('The device step does not shift the particles (PairSums.shiftingDisplacement gives the '
                             'shift as a host step); the delta-SPH scheme sets particleShifting to False.')
        self._config = config
        self._boundary = boundary
        self._particles: ParticleSystem | None = None
        self._time, self._step, self._dt = 0.0, 0, config.maxTimeStep
        self._gravity = abs(float(np.asarray(config.gravity)[-1]))
        self._speedOfSound = (float(config.speedOfSound) if config.speedOfSound is not None
                              else speedOfSound(float(config.domainSize[-1]), config.speedOfSoundFactor, self._gravity))
        gpu = PairSums(config.dimensions, config.smoothingLength, config.domainMin, config.domainMax,
                       config.periodicWidth or 0.0, kernelType = config.kernelType,
                       pairPrecision = config.pairPrecision)
        gpu.setDiffusion(config.densityDiffusion, config.diffusionDelta, self._speedOfSound, config.referenceDensity,
                         self._gravity)
        gpu.setShifting(config.particleShifting, config.shiftingMach, config.particleSpacing, self._speedOfSound,
                        config.shiftingSpeed == 'local')
        gpu.setWallPressure(config.wallPressure == 'extrapolated')
        gpu.setNegativePressure(config.negativePressure)
        gpu.setContinuityCorrection(config.continuityCorrection)
        gpu.setViscosityLimiter(config.viscosityLimiter)
        gpu.setClosure(config.turbulenceClosure == 'deltaLes', self._speedOfSound, config.referenceDensity, 1.0034e-6)
        gpu.setFilterInterval(config.densityFilterInterval)
        self._gpu = gpu
        self._energy: dict | None = None

    def _pressure(self, densities, isFluid) -> np.ndarray:
        config = self._config
        pressures = taitEquationOfState(densities, config.referenceDensity, self._speedOfSound,
                                        clampNegative = not config.negativePressure)
        if config.negativePressure:
            pressures[~isFluid] = np.maximum(pressures[~isFluid], 0.0)
        return pressures

    def initialize(self, particles: ParticleSystem) -> None:
        '''Take the particles and set their starting density hydrostatic beneath the initial surface (sphSolver.py:1869-1897).'''
        self._particles = particles
        self._time, self._step = 0.0, 0
        surface = self._config.initialSurface
        if surface is None:
            particles.densities = np.full(particles.nParticles, self._config.referenceDensity)
        else:
            x = particles.positions[:, 0]
            elevation = surface(x) if callable(surface) else np.full_like(x, float(surface))
            particles.densities = hydrostaticDensity(elevation - particles.positions[:, -1],
                                                     self._config.referenceDensity, self._speedOfSound, self._gravity)
        particles.pressures = self._pressure(particles.densities, particles.isFluid)
        self._gpu.update(particles.positions, particles.masses, particles.isFluid, particles.lengths)

    def _prepare(self) -> None:
        config, gpu, p = self._config, self._gpu, self._particles
        gpu.setPhysics(config.referenceDensity, self._speedOfSound, config.alphaViscosity,
                       config.boundaryViscosityFactor, np.asarray(config.gravity, dtype = float), config.xsphEpsilon,
                       config.densityFloor * config.referenceDensity, config.cflNumber, config.maxTimeStep)
        gpu.push(p)
        gpu.setBoundary(self._boundary, p)

    def step(self) -> SimulationState:
        '''One step on the device, the time step from the host (sphSolver.py:1899-1939 with 1976-1988).'''
        self._prepare()
        gpu, config = self._gpu, self._config
        interval = config.densityFilterInterval
        maxSpeed, maxAcceleration = gpu.residentRates(bool(interval) and self._step > 0 and self._step % interval == 0)
        self._dt = cflTimeStep(self.shortestLength, self._speedOfSound, maxSpeed, maxAcceleration, config.cflNumber,
                               config.maxTimeStep)
        gpu.residentIntegrate(self._dt)
        self._time += self._dt
        self._step += 1
        self._collectEnergy()
        gpu.pull(self._particles)
        return self.currentState

    def _collectEnergy(self) -> None:
        '''Add the device ledger's totals since the last push to the budget (sphSolver.py:1987-1989).'''
        if self._energy is None:
            return
        from pysurf_amd._native import energyTermNames
        for name, value in zip(energyTermNames, self._gpu.energySums()):
            self._energy[name] += float(value)

    def trackEnergy(self, enabled: bool = True) -> None:
        '''Keep the energy budget from here on (sphSolver.py:2128-2149): every term starts at zero.'''
        from pysurf_amd._native import energyTermNames
        self._energy = {name: 0.0 for name in energyTermNames + ['refinement']} if enabled else None
        self._gpu.setEnergyBudget(enabled)

    @property
    def energyBudget(self) -> dict | None:
        '''Each term's energy since trackEnergy [J], or None while the budget is off (sphSolver.py:2151-2154).'''
        return None if self._energy is None else dict(self._energy)

    @property
    def elasticEnergy(self) -> float:
        '''The fluid's elastic energy, sum m e(rho) [J] (sphSolver.py:2156-2161).'''
        p = self._particles
        fluid = p.isFluid
        energy = elasticEnergy(p.densities[fluid], self._config.referenceDensity, self._speedOfSound,
                               self._config.negativePressure)
        return float(np.sum(p.masses[fluid] * energy))

    def advance(self, untilTime: float, maxSteps: int | None = None, chunk: int = 64) -> SimulationState:

        '''

        Step until the simulated time reaches `untilTime`, or `maxSteps` steps, every step on the
        device with the time step reduced there (sphSolver.py:1941-1992).

        Raises:
        -------
        NumericalInstabilityError : If the state stops being finite.

        '''

        lastStep = self._step + maxSteps if maxSteps is not None else 2**62
        if not (self._time < untilTime and self._step < lastStep):
            return self.currentState
        self._prepare()
        clock = self._gpu.advance(self._time, untilTime, self._step, lastStep, chunk = chunk)
        self._time, self._step, self._dt = clock['time'], clock['steps'], clock['dt']
        self._collectEnergy()
        self._gpu.pull(self._particles)
        if not np.all(np.isfinite(self._particles.positions)):
            raise NumericalInstabilityError(f'Particle positions stopped being finite by t = {self._time:.6f} s.')
        return self.currentState

    def sortParticles(self) -> np.ndarray:
        '''Reorder the fluid among its slots in cell order; the boundary keeps its slots (sphSolver.py:2093-2126).'''
        p = self._particles
        fluid = np.flatnonzero(p.isFluid)
        key = cellOrderKey(p.positions[fluid], 2.0 * self._gpu.smoothingLength, self._gpu.gridShape)
        order = np.arange(p.nParticles)
        order[fluid] = fluid[np.argsort(key, kind = 'stable')]
        p.permute(order)
        return order

    def restore(self, particles: ParticleSystem, time: float, step: int, timeStep: float) -> None:
        '''Take up a saved state in place of `initialize` (sphSolver.py:2163-2196).'''
        self._particles = particles
        self._time, self._step, self._dt = float(time), int(step), float(timeStep)
        self._gpu.update(particles.positions, particles.masses, particles.isFluid, particles.lengths)

    @property
    def clock(self) -> dict:
        return {'time': self._time, 'step': self._step, 'timeStep': self._dt}

    @property
    def shortestLength(self) -> float:
        p = self._particles
        if p is None or p.lengths is None:
            return self._config.smoothingLength
        return float(p.lengths.min())

    @property
    def currentState(self) -> SimulationState:
        p = self._particles
        return SimulationState(time = self._time, step = self._step, dt = self._dt, kineticEnergy = p.kineticEnergy(),
                               potentialEnergy = p.potentialEnergy(self._gravity), maxVelocity = p.maxSpeed(),
                               maxDensityError = p.maxDensityError(self._config.referenceDensity))

    @property
    def particles(self) -> ParticleSystem:
        return self._particles

    @property
    def config(self) -> SimulationConfig:
        return self._config

    @property
    def time(self) -> float:
        return self._time

    @property
    def speedOfSound(self) -> float:
        return self._speedOfSound

    @property
    def eosStiffness(self) -> float:
        return self._config.referenceDensity * self._speedOfSound**2 / taitExponent
