'''

Particle splitting and merging for a refined zone, after pySurf's `sphRefinement`
(src/pysurf/sphRefinement.py:101-311): a split turns each chosen fluid particle into 2^dim daughters
at the corners of a square (cube) half its spacing across, each with 1/2^dim of its mass and its
smoothing length times `lengthRatio`; a merge turns groups of 2^dim split particles of one level,
binned into cells the size of a particle one level up, back into one.

These are the host's forms; BatchEngine.split and BatchEngine.merge do the same on the device.

'''

from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np


@dataclass
class MergeResult:
    '''What a merge did (sphRefinement.py:101-121): the groups merged, the kinetic energy of the members' motion about their groups' centres of mass it removed, and each remaining particle's index before.'''

    groups: int = 0
    removedEnergy: float = 0.0
    source: np.ndarray = field(default_factory = lambda: np.zeros(0, dtype = np.intp))


def refinementLevels(particles, referenceMass: float) -> np.ndarray:
    '''How many times each particle has split: log2(referenceMass / m) / dim, rounded (sphRefinement.py:123-133).'''
    ratio = referenceMass / np.asarray(particles.masses, dtype = float)
    return np.rint(np.log2(ratio) / particles.dimensions).astype(int)


def splitParticles(particles, mask: np.ndarray, referenceDensity: float, referenceLength: float,
                   lengthRatio: float = 1.0) -> int:

    '''

    Split the fluid particles `mask` selects, in place; the first daughter takes its mother's slot
    and the others are appended, mother by mother. Returns how many particles split.

    Parameters:
    -----------
    particles : ParticleSystem
        The state; its arrays are replaced by longer ones
    mask : np.ndarray
        Particles to split, shape (N,); boundary particles in it stay whole
    referenceDensity : float
        rho0: a particle's spacing is (m / rho0)^(1/dim)
    referenceLength : float
        The smoothing length of a particle that has not split
    lengthRatio : float
        A daughter's smoothing length over its mother's

    '''

    chosen = np.flatnonzero(np.asarray(mask, dtype = bool) & particles.isFluid)
    if chosen.size == 0:
        return 0
    dimensions = particles.dimensions
    count = 2**dimensions
    lengths = (particles.lengths.astype(float, copy = True) if particles.lengths is not None
               else np.full(particles.nParticles, float(referenceLength)))
    corners = np.array(list(itertools.product((-0.25, 0.25), repeat = dimensions)))
    spacing = (particles.masses[chosen] / referenceDensity)**(1.0 / dimensions)
    daughters = particles.positions[chosen][:, None, :] + spacing[:, None, None] * corners[None]
    mass = particles.masses[chosen] / count
    length = lengths[chosen] * lengthRatio
    particles.positions[chosen] = daughters[:, 0]
    particles.masses[chosen] = mass
    lengths[chosen] = length

    def repeat(values):
        return np.repeat(values, count - 1, axis = 0)

    particles.positions = np.vstack([particles.positions, daughters[:, 1:].reshape(-1, dimensions)])
    particles.velocities = np.vstack([particles.velocities, repeat(particles.velocities[chosen])])
    particles.accelerations = np.vstack([particles.accelerations, repeat(particles.accelerations[chosen])])
    particles.densities = np.concatenate([particles.densities, repeat(particles.densities[chosen])])
    particles.pressures = np.concatenate([particles.pressures, repeat(particles.pressures[chosen])])
    particles.masses = np.concatenate([particles.masses, repeat(mass)])
    particles.isFluid = np.concatenate([particles.isFluid, np.ones(chosen.size * (count - 1), dtype = bool)])
    particles.lengths = np.concatenate([lengths, repeat(length)])
    if np.all(particles.lengths == referenceLength):
        particles.lengths = None
    return int(chosen.size)


def cellGroups(positions: np.ndarray, cell: float, offset: float, size: int) -> np.ndarray:

    '''

    Groups of `size` points binned into cells `cell` across, the grid shifted by `offset` cells: in
    each cell that holds at least `size`, the `size` nearest its centre, ties by index
    (sphRefinement.py:188-210). Returns the points' indices, (groups, size), cells in row-major
    order of their integer coordinates.

    '''

    scaled = positions / cell + offset
    keys = np.floor(scaled).astype(np.int64)
    keys -= keys.min(axis = 0)
    flat = np.ravel_multi_index(keys.T, keys.max(axis = 0) + 1)
    distance = np.sum((scaled - np.floor(scaled) - 0.5)**2, axis = 1)
    order = np.lexsort((distance, flat))
    cells = flat[order]
    starts = np.flatnonzero(np.r_[True, cells[1:] != cells[:-1]])
    counts = np.diff(np.r_[starts, len(order)])
    owner = np.repeat(np.arange(len(starts)), counts)
    rank = np.arange(len(order)) - starts[owner]
    taken = (rank < size) & (counts[owner] >= size)
    return order[taken].reshape(-1, size)


def mergeParticles(particles, mask: np.ndarray, referenceDensity: float, referenceMass: float, referenceLength: float,
                   lengthRatio: float = 1.0, offsets: tuple = (0.0, 0.5)) -> MergeResult:

    '''

    Merge the split fluid particles `mask` selects, in place, 2^dim of one level at a time into one
    particle of the level above (sphRefinement.py:212-311). Level by level, a pass per grid shift in
    `offsets` groups the level's free particles by cellGroups in cells the size of a particle one
    level up. A group becomes one particle in its first member's slot: the mass-weighted mean
    position, velocity, acceleration and pressure, the total mass, the density of the total mass
    over the members' volumes, and the members' mean length over `lengthRatio`. The other members
    are removed; fluid particles from the end fill their places among the boundary's, so the
    boundary keeps its indices.

    '''

    dimensions = particles.dimensions
    size = 2**dimensions
    levels = refinementLevels(particles, referenceMass)
    eligible = np.asarray(mask, dtype = bool) & particles.isFluid & (levels > 0)
    groups = []
    for level in np.unique(levels[eligible]):
        cell = (referenceMass / size**(level - 1) / referenceDensity)**(1.0 / dimensions)
        free = np.flatnonzero(eligible & (levels == level))
        for offset in offsets:
            if free.size < size:
                break
            members = free[cellGroups(particles.positions[free], cell, offset, size)]
            if members.size == 0:
                continue
            groups.append(members)
            free = np.setdiff1d(free, members.ravel(), assume_unique = True)
    if not groups:
        return MergeResult(source = np.arange(particles.nParticles))
    members = np.sort(np.vstack(groups), axis = 1)
    lengths = (particles.lengths.astype(float, copy = True) if particles.lengths is not None
               else np.full(particles.nParticles, float(referenceLength)))
    m = particles.masses[members]
    total = m.sum(axis = 1)
    weights = m / total[:, None]
    velocity = np.einsum('gk,gkd->gd', weights, particles.velocities[members])
    relative = particles.velocities[members] - velocity[:, None, :]
    removed = 0.5 * float(np.sum(m * np.sum(relative**2, axis = 2)))
    first = members[:, 0]
    particles.positions[first] = np.einsum('gk,gkd->gd', weights, particles.positions[members])
    particles.accelerations[first] = np.einsum('gk,gkd->gd', weights, particles.accelerations[members])
    particles.pressures[first] = np.sum(weights * particles.pressures[members], axis = 1)
    particles.densities[first] = total / np.sum(m / particles.densities[members], axis = 1)
    particles.velocities[first] = velocity
    particles.masses[first] = total
    lengths[first] = lengths[members].mean(axis = 1) / lengthRatio
    particles.lengths = lengths
    source = compactionSource(particles.nParticles, np.sort(members[:, 1:].ravel()), particles.isFluid)
    for name in ('positions', 'velocities', 'accelerations', 'densities', 'pressures', 'masses', 'isFluid', 'lengths'):
        setattr(particles, name, getattr(particles, name)[source])
    return MergeResult(groups = len(members), removedEnergy = removed, source = source)


def compactionSource(count: int, removed: np.ndarray, isFluid: np.ndarray) -> np.ndarray:

    '''

    Each remaining particle's index before `removed` (sorted) leave (sphRefinement.py:291-306):
    the holes at or before the last boundary particle take, in order, the last as many remaining
    particles after it; everything else keeps its order. Raises RuntimeError when too few follow.

    '''

    alive = np.ones(count, dtype = bool)
    alive[removed] = False
    boundary = np.flatnonzero(~np.asarray(isFluid, dtype = bool))
    last = int(boundary[-1]) if boundary.size else -1
    holes = removed[removed <= last]
    tail = np.flatnonzero(alive[last + 1:]) + last + 1
    if tail.size < holes.size:
        raise RuntimeError('Too few fluid particles follow the boundary to fill the places a merge leaves among it.')
    movers = tail[tail.size - holes.size:]
    source = np.arange(count)
    source[holes] = movers
    alive[holes] = True
    alive[movers] = False
    return source[alive]
