'''

The host's particle state and the lattice seeding the scenario starts are laid out with.

`ParticleSystem` keeps the field names and shapes of pySurf's container (src/pysurf/sphParticles.py:
41-283), which the drop-in interface reads and writes in place: vectors (N, dim), scalars (N,),
masses per unit depth in 2D, fluid first and boundary after. Its operations are written over the
list of its array fields, so copying, permuting and counting treat every field alike.

Seeding: a square lattice of cell centres, a predicate's share of it, and a seeded jitter that
moves each point off the lattice while keeping it inside the water.

'''

from __future__ import annotations

from dataclasses import dataclass, fields
from typing import Callable

import numpy as np


@dataclass
class ParticleSystem:

    '''

    Particle state, structure of arrays (sphParticles.py:41-124). `lengths` is None while every
    particle has the run's smoothing length, and an (N,) array once some have split.

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

        '''

        The state of a start laid on a lattice of `spacing`: the fluid's points, then the
        boundary's, every particle carrying its lattice cell's mass at rest density, rho0 s^dim,
        and starting at rest density with no pressure; the fluid moves with `fluidVelocities` (at
        rest without them), the boundary stays at rest (sphParticles.py:81-124).

        '''

        fluid = np.asarray(fluidPositions, dtype = float)
        dimensions = fluid.shape[1]
        walls = np.asarray(boundaryPositions, dtype = float).reshape(-1, dimensions)
        kinds = np.zeros(len(fluid) + len(walls), dtype = bool)
        kinds[:len(fluid)] = True
        count = len(kinds)
        moving = np.zeros((count, dimensions))
        if fluidVelocities is not None:
            moving[:len(fluid)] = fluidVelocities
        return cls(positions = np.concatenate([fluid, walls]), velocities = moving,
                   accelerations = np.zeros((count, dimensions)), densities = np.full(count, referenceDensity),
                   pressures = np.zeros(count), masses = np.full(count, referenceDensity * spacing**dimensions),
                   isFluid = kinds)

    def _arrays(self):
        '''(name, array) of every field that holds one.'''
        for entry in fields(self):
            value = getattr(self, entry.name)
            if value is not None:
                yield entry.name, value

    def copy(self) -> ParticleSystem:
        '''An independent copy of every field.'''
        return ParticleSystem(**{entry.name: None for entry in fields(self)} | {name: values.copy()
                                                                              for name, values in self._arrays()})

    def permute(self, order: np.ndarray) -> None:
        '''Reorder every field: slot k takes what slot order[k] held (sphParticles.py:186-203).'''
        for name, values in list(self._arrays()):
            setattr(self, name, values[order])

    @property
    def nParticles(self) -> int:
        return len(self.positions)

    @property
    def nFluid(self) -> int:
        return int(np.count_nonzero(self.isFluid))

    @property
    def dimensions(self) -> int:
        return self.positions.shape[1]

    def kineticEnergy(self) -> float:
        '''The fluid's kinetic energy, sum m |v|^2 / 2 (sphParticles.py:236-242).'''
        speed2 = np.einsum('ij,ij->i', self.velocities, self.velocities)
        return float(0.5 * np.dot(self.masses[self.isFluid], speed2[self.isFluid]))

    def potentialEnergy(self, gravity: float) -> float:
        '''The fluid's potential energy, sum m g z with g the downward gravity (sphParticles.py:244-255).'''
        return float(gravity * np.dot(self.masses[self.isFluid], self.positions[self.isFluid, -1]))

    def maxSpeed(self) -> float:
        '''The fluid's largest speed, 0 with no fluid (sphParticles.py:257-265).'''
        speeds = np.linalg.norm(self.velocities[self.isFluid], axis = 1)
        return float(speeds.max()) if speeds.size else 0.0

    def maxDensityError(self, referenceDensity: float) -> float:
        '''The fluid's largest density departure from rho0, over rho0 (sphParticles.py:267-280).'''
        departures = np.abs(self.densities[self.isFluid] - referenceDensity)
        return float(departures.max()) / referenceDensity if departures.size else 0.0


def latticePoints(lower, upper, spacing: float) -> np.ndarray:
    '''The centres of the lattice cells of `spacing` that fill the box [lower, upper], x fastest (sphParticles.py:289-312).'''
    centres = [np.arange(low + spacing / 2.0, high, spacing) for low, high in zip(lower, upper)]
    return np.stack([axis.ravel() for axis in np.meshgrid(*centres, indexing = 'xy')], axis = 1)


def fillRegion(inside: Callable[[np.ndarray], np.ndarray], lower, upper, spacing: float) -> np.ndarray:
    '''The lattice points of the box [lower, upper] that `inside` accepts (sphParticles.py:425-448).'''
    points = latticePoints(lower, upper, spacing)
    return points[np.asarray(inside(points), dtype = bool)]


def _mirror(values: np.ndarray, low, high) -> np.ndarray:
    '''Values folded back into [low, high] by one reflection at each end: below low to 2 low - v, above high to 2 high - v.'''
    folded = np.where(values < low, 2.0 * low - values, values)
    return np.where(folded > high, 2.0 * high - folded, folded)


def jitterPoints(points: np.ndarray, spacing: float, amplitude: float, seed: int, lower, upper,
                 floor: Callable[..., np.ndarray], ceiling: Callable[..., np.ndarray] | None = None) -> np.ndarray:

    '''

    The points moved off their lattice by a seeded uniform offset of up to `amplitude` spacings on
    every axis, and folded back into the water (sphParticles.py:357-423 gives the rules).

    The water a point may occupy: along each horizontal axis between the walls half a spacing in
    from `lower` and `upper`; vertically from half a spacing above `floor` (a function of the
    horizontal coordinates, one array each) up to, and not including, `ceiling` (likewise; None for
    no surface). A moved point is mirrored once at each face it crossed, horizontally first and then
    vertically at the floor and ceiling under its new horizontal place. Where water is thinner than
    the offset a point can still end outside; such a point keeps its lattice place. No point is
    lost, so the water's volume is the lattice's.

    '''

    start = np.asarray(points, dtype = float)
    if amplitude <= 0.0:
        return start.copy()
    moved = start + np.random.default_rng(seed).uniform(-amplitude, amplitude, start.shape) * spacing
    across = start.shape[1] - 1
    lower, upper = np.asarray(lower, dtype = float), np.asarray(upper, dtype = float)
    for axis in range(across):
        moved[:, axis] = _mirror(moved[:, axis], lower[axis] + 0.5 * spacing, upper[axis] - 0.5 * spacing)
    horizontal = moved[:, :across].T
    ground = np.asarray(floor(*horizontal), dtype = float) + 0.5 * spacing
    surface = np.full(len(moved), np.inf) if ceiling is None else np.asarray(ceiling(*horizontal), dtype = float)
    moved[:, -1] = _mirror(moved[:, -1], ground, surface - 1e-9 * spacing)
    # what the mirrors could not bring back in, within a billionth of a spacing
    margin = (0.5 - 1e-9) * spacing
    between = np.all((moved[:, :across] >= lower + margin) & (moved[:, :across] <= upper - margin), axis = 1)
    outside = ~between | (moved[:, -1] < ground - 1e-9 * spacing) | (moved[:, -1] >= surface)
    moved[outside] = start[outside]
    return moved
