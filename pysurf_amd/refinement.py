'''

Splitting and merging particles in a refined zone, on the host (the rules: pySurf's
src/pysurf/sphRefinement.py:101-311). BatchEngine.split and BatchEngine.merge do the same on the
device; the host forms build the refined starts and check the device's.

A split replaces a fluid particle by 2^dim daughters on the corners of a square (cube) half its
spacing across, each with 1/2^dim of its mass and its smoothing length times `lengthRatio`. A merge
turns 2^dim particles of one level that share a cell the size of their mother back into one. Both
are written over the state's list of per-particle arrays: a split appends rows copied from their
mothers, a merge rebuilds the rows through one gather that keeps the boundary's indices.

'''

from __future__ import annotations

from dataclasses import dataclass, field, fields

import numpy as np


@dataclass
class MergeResult:
    '''What a merge did (sphRefinement.py:101-121): how many groups merged, the kinetic energy of the members' motion about their groups' centres of mass it took away, and for each remaining particle its index before.'''

    groups: int = 0
    removedEnergy: float = 0.0
    source: np.ndarray = field(default_factory = lambda: np.zeros(0, dtype = np.intp))


def refinementLevels(particles, referenceMass: float) -> np.ndarray:
    '''How many times each particle has split, the base-2^dim logarithm of referenceMass / m rounded (sphRefinement.py:123-133).'''
    return np.rint(np.log2(referenceMass / np.asarray(particles.masses, dtype = float)) / particles.dimensions).astype(int)


def _daughterCorners(dimensions: int) -> np.ndarray:
    '''The 2^dim corner offsets in spacings, daughter k's axis a at +1/4 where bit (dim - 1 - a) of k is set and -1/4 where it is clear.'''
    bits = (np.arange(2**dimensions)[:, None] >> np.arange(dimensions - 1, -1, -1)[None, :]) & 1
    return (bits - 0.5) * 0.5


def _lengthsOf(particles, referenceLength: float) -> np.ndarray:
    '''Every particle's smoothing length as a fresh array.'''
    if particles.lengths is None:
        return np.full(particles.nParticles, float(referenceLength))
    return particles.lengths.astype(float, copy = True)


def _rows(particles):
    '''(name, array) of the state's per-particle arrays that are set.'''
    for entry in fields(particles):
        values = getattr(particles, entry.name)
        if values is not None:
            yield entry.name, values


def splitParticles(particles, mask: np.ndarray, referenceDensity: float, referenceLength: float,
                   lengthRatio: float = 1.0) -> int:

    '''

    Split the fluid particles `mask` selects, in place (sphRefinement.py:135-186). Each mother's
    slot takes its first daughter (the corner at -1/4 on every axis); the other daughters are
    appended mother by mother, in corner order, carrying their mother's velocity, acceleration,
    density and pressure. A daughter's spacing is half its mother's, (m / rho0)^(1/dim) / 2, with
    rho0 `referenceDensity`; `referenceLength` is the length of a particle that never split.
    Boundary particles in `mask` stay whole. Returns how many particles split.

    '''

    mothers = np.flatnonzero(np.asarray(mask, dtype = bool) & particles.isFluid)
    if mothers.size == 0:
        return 0
    dimensions = particles.dimensions
    corners = _daughterCorners(dimensions)
    spacing = (particles.masses[mothers] / referenceDensity)**(1.0 / dimensions)
    places = particles.positions[mothers][:, None, :] + spacing[:, None, None] * corners[None]
    particles.lengths = _lengthsOf(particles, referenceLength)
    particles.lengths[mothers] = particles.lengths[mothers] * lengthRatio
    particles.masses[mothers] = particles.masses[mothers] / len(corners)
    particles.positions[mothers] = places[:, 0]
    # the appended daughters are copies of their (already divided) mothers but for their places
    copies = np.repeat(mothers, len(corners) - 1)
    for name, values in list(_rows(particles)):
        setattr(particles, name, np.concatenate([values, values[copies]]))
    particles.positions[-copies.size:] = places[:, 1:].reshape(-1, dimensions)
    if np.all(particles.lengths == referenceLength):
        particles.lengths = None
    return int(mothers.size)


def _groupsInCells(positions: np.ndarray, cell: float, offset: float, size: int) -> np.ndarray:

    '''

    Groups of `size` points sharing a cell of a grid `cell` across, shifted by `offset` cells
    (sphRefinement.py:188-210): every cell holding `size` or more gives one group, its `size`
    points nearest the cell's centre, equal distances by index. Returns the points' indices, shape
    (groups, size), the cells in row-major order of their integer coordinates.

    '''

    scaled = positions / cell + offset
    corner = np.floor(scaled)
    inCell = scaled - corner - 0.5
    distance = np.sum(inCell**2, axis = 1)
    cells, owner, population = np.unique(corner.astype(np.int64), axis = 0, return_inverse = True,
                                         return_counts = True)
    owner = owner.reshape(-1)
    order = np.lexsort((distance, owner))
    ranked = owner[order]
    rank = np.arange(order.size) - np.searchsorted(ranked, ranked, side = 'left')
    chosen = (rank < size) & (population[ranked] >= size)
    return order[chosen].reshape(-1, size)


def _mergeGroups(particles, eligible: np.ndarray, levels: np.ndarray, referenceDensity: float,
                 referenceMass: float, offsets) -> np.ndarray | None:

    '''

    The groups a merge takes, shape (groups, 2^dim), each row ascending; None when there are none.
    Level by level (coarsest first), the level's free particles are grouped in cells the size of a
    particle one level up, once per grid shift in `offsets`; a particle grouped by one shift is
    no longer free for the next.

    '''

    dimensions = particles.dimensions
    size = 2**dimensions
    found = []
    for level in np.unique(levels[eligible]):
        cell = (referenceMass / size**(level - 1) / referenceDensity)**(1.0 / dimensions)
        free = np.flatnonzero(eligible & (levels == level))
        for offset in offsets:
            if free.size < size:
                break
            groups = free[_groupsInCells(particles.positions[free], cell, offset, size)]
            if groups.size:
                found.append(groups)
                free = np.setdiff1d(free, groups.ravel(), assume_unique = True)
    return np.sort(np.concatenate(found), axis = 1) if found else None


def mergeParticles(particles, mask: np.ndarray, referenceDensity: float, referenceMass: float, referenceLength: float,
                   lengthRatio: float = 1.0, offsets: tuple = (0.0, 0.5)) -> MergeResult:

    '''

    Merge the split fluid particles `mask` selects, in place, 2^dim of one level at a time into one
    particle of the level above (sphRefinement.py:212-311). The merged particle takes its group's
    lowest slot: the total mass; the mass-weighted mean position, velocity, acceleration and
    pressure; the density of the total mass over the members' volumes sum m / rho; the members'
    mean smoothing length over `lengthRatio`. The members' kinetic energy about the group's mean
    velocity is what the merge takes away. The other members' slots are closed (compactionSource).

    '''

    levels = refinementLevels(particles, referenceMass)
    eligible = np.asarray(mask, dtype = bool) & particles.isFluid & (levels > 0)
    groups = _mergeGroups(particles, eligible, levels, referenceDensity, referenceMass, offsets)
    if groups is None:
        return MergeResult(source = np.arange(particles.nParticles))
    masses = particles.masses[groups]
    total = masses.sum(axis = 1)
    weights = masses / total[:, None]

    def weighted(values):
        return np.einsum('gk,gkd->gd', weights, values[groups])

    mean = weighted(particles.velocities)
    spread = particles.velocities[groups] - mean[:, None, :]
    removed = 0.5 * float(np.sum(masses * np.sum(spread**2, axis = 2)))
    kept = groups[:, 0]
    lengths = _lengthsOf(particles, referenceLength)
    lengths[kept] = lengths[groups].mean(axis = 1) / lengthRatio
    particles.lengths = lengths
    particles.positions[kept] = weighted(particles.positions)
    particles.accelerations[kept] = weighted(particles.accelerations)
    particles.pressures[kept] = np.sum(weights * particles.pressures[groups], axis = 1)
    particles.densities[kept] = total / np.sum(masses / particles.densities[groups], axis = 1)
    particles.velocities[kept] = mean
    particles.masses[kept] = total
    source = compactionSource(particles.nParticles, np.sort(groups[:, 1:].ravel()), particles.isFluid)
    particles.permute(source)
    return MergeResult(groups = len(groups), removedEnergy = removed, source = source)


def compactionSource(count: int, removed: np.ndarray, isFluid: np.ndarray) -> np.ndarray:

    '''

    For each particle left after the sorted slots `removed` close, its slot before
    (sphRefinement.py:291-306). Up to the last boundary particle every slot stays, a closed one
    taking, in order, the last as many remaining particles after it; after the last boundary
    particle the rest follow in order. Raises RuntimeError when too few remain after it.

    '''

    kinds = np.asarray(isFluid, dtype = bool)
    walls = np.flatnonzero(~kinds)
    last = int(walls[-1]) if walls.size else -1
    open_ = np.ones(count, dtype = bool)
    open_[removed] = False
    holes = removed[removed <= last]
    after = np.flatnonzero(open_[last + 1:]) + last + 1
    if after.size < holes.size:
        raise RuntimeError('Too few fluid particles follow the boundary to fill the places a merge leaves among it.')
    head = np.arange(last + 1)
    head[holes] = after[after.size - holes.size:]
    return np.concatenate([head, after[:after.size - holes.size]])
