'''

Particle state and lattice seeding: the SoA container the engine pushes and pulls, after
pySurf's `sphParticles` (src/pysurf/sphParticles.py:41-448).

'''

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable

import numpy as np


@dataclass
class ParticleSystem:

    '''

    SPH particle state (sphParticles.py:41-124): vectors shaped (N, dim), scalars (N,); masses
    per unit depth in 2D.

    '''

    positions: np.ndarray
    velocities: np.ndarray
    accelerations: np.ndarray
    densities: np.ndarray
    pressures: np.ndarray
    masses: np.ndarray
    isFluid: np.ndarray
    lengths: np.ndarray | None = None

    @classmethod
    def assemble(cls, fluidPositions: np.ndarray, boundaryPositions: np.ndarray, spacing: float,
                 referenceDensity: float, fluidVelocities: np.ndarray | None = None) -> ParticleSystem:
        '''Fluid then boundary particles, each of mass rho0 s^dim, at rest density (sphParticles.py:81-124).'''
        fluidPositions = np.asarray(fluidPositions, dtype = float)
        boundaryPositions = np.asarray(boundaryPositions, dtype = float).reshape(-1, fluidPositions.shape[1])
        nFluid, nBoundary = len(fluidPositions), len(boundaryPositions)
        nTotal, dimensions = nFluid + nBoundary, fluidPositions.shape[1]
        velocities = np.zeros((nTotal, dimensions))
        if fluidVelocities is not None:
            velocities[:nFluid] = fluidVelocities
        return cls(positions = np.vstack([fluidPositions, boundaryPositions]), velocities = velocities,
                   accelerations = np.zeros((nTotal, dimensions)), densities = np.full(nTotal, referenceDensity),
                   pressures = np.zeros(nTotal), masses = np.full(nTotal, referenceDensity * spacing**dimensions),
                   isFluid = np.concatenate([np.ones(nFluid, dtype = bool), np.zeros(nBoundary, dtype = bool)]))

    def copy(self) -> ParticleSystem:
        return ParticleSystem(**{name: (None if getattr(self, name) is None else getattr(self, name).copy())
                                 for name in ('positions', 'velocities', 'accelerations', 'densities', 'pressures',
                                              'masses', 'isFluid', 'lengths')})

    def permute(self, order: np.ndarray) -> None:
        '''Slot k takes what slot order[k] held (sphParticles.py:186-203).'''
        for name in ('positions', 'velocities', 'accelerations', 'densities', 'pressures', 'masses', 'isFluid'):
            setattr(self, name, getattr(self, name)[order])
        if self.lengths is not None:
            self.lengths = self.lengths[order]

    @property
    def nParticles(self) -> int:
        return self.positions.shape[0]

    @property
    def nFluid(self) -> int:
        return int(np.sum(self.isFluid))

    @property
    def dimensions(self) -> int:
        return self.positions.shape[1]

    def kineticEnergy(self) -> float:
        '''sum of m |v|^2 / 2 over the fluid (sphParticles.py:236-242).'''
        fluid = self.isFluid
        return float(0.5 * np.sum(self.masses[fluid] * np.sum(self.velocities[fluid]**2, axis = 1)))

    def potentialEnergy(self, gravity: float) -> float:
        '''sum of m g z over the fluid (sphParticles.py:244-255).'''
        fluid = self.isFluid
        return float(np.sum(self.masses[fluid] * gravity * self.positions[fluid, -1]))

    def maxSpeed(self) -> float:
        '''Largest fluid speed (sphParticles.py:257-265).'''
        fluid = self.isFluid
        if not np.any(fluid):
            return 0.0
        return float(np.max(np.linalg.norm(self.velocities[fluid], axis = 1)))

    def maxDensityError(self, referenceDensity: float) -> float:
        '''Largest relative fluid density departure (sphParticles.py:267-280).'''
        fluid = self.isFluid
        if not np.any(fluid):
            return 0.0
        return float(np.max(np.abs(self.densities[fluid] - referenceDensity)) / referenceDensity)


def latticePoints(lower, upper, spacing: float) -> np.ndarray:
    '''Cell centres of a square lattice in [lower, upper], half a spacing in from each face (sphParticles.py:289-312).'''
    axes = [np.arange(low + spacing / 2.0, high, spacing) for low, high in zip(lower, upper)]
    grids = np.meshgrid(*axes, indexing = 'xy')
    return np.column_stack([grid.ravel() for grid in grids])


def fillRegion(inside: Callable[[np.ndarray], np.ndarray], lower, upper, spacing: float) -> np.ndarray:
    '''The lattice points of a box that a predicate accepts (sphParticles.py:425-448).'''
    points = latticePoints(lower, upper, spacing)
    return points[np.asarray(inside(points), dtype = bool)]


def jitterPoints(points: np.ndarray, spacing: float, amplitude: float, seed: int, lower, upper,
                 floor: Callable[..., np.ndarray], ceiling: Callable[..., np.ndarray] | None = None) -> np.ndarray:
    '''

    Points moved by a seeded uniform offset of up to `amplitude` spacings per axis, reflected back
    across any face they crossed; a point the reflections leave outside keeps its place
    (sphParticles.py:357-423).

    '''

    points = np.asarray(points, dtype = float)
    if amplitude <= 0.0:
        return points.copy()
    moved = points + np.random.default_rng(seed).uniform(-amplitude, amplitude, points.shape) * spacing
    horizontal = points.shape[1] - 1
    lower, upper = np.asarray(lower, dtype = float), np.asarray(upper, dtype = float)

    def reflect(values, low, high):
        values = np.where(values < low, 2.0 * low - values, values)
        return np.where(values > high, 2.0 * high - values, values)

    for axis in range(horizontal):
        moved[:, axis] = reflect(moved[:, axis], lower[axis] + 0.5 * spacing, upper[axis] - 0.5 * spacing)
    bottom = np.asarray(floor(*moved[:, :horizontal].T), dtype = float) + 0.5 * spacing
    top = (np.asarray(ceiling(*moved[:, :horizontal].T), dtype = float) if ceiling is not None
           else np.full(len(moved), np.inf))
    moved[:, -1] = reflect(moved[:, -1], bottom, top - 1e-9 * spacing)
    walls = np.all((moved[:, :horizontal] >= lower + (0.5 - 1e-9) * spacing)
                   & (moved[:, :horizontal] <= upper - (0.5 - 1e-9) * spacing), axis = 1)
    stray = ~walls | (moved[:, -1] < bottom - 1e-9 * spacing) | (moved[:, -1] >= top)
    moved[stray] = points[stray]
    return moved
