'''

Particle splitting for a refined zone, after pySurf's `sphRefinement.splitParticles`
(src/pysurf/sphRefinement.py:136-190): each chosen fluid particle becomes 2^dim daughters at the
corners of a square (cube) half its spacing across, each with 1/2^dim of its mass and its smoothing
length times `lengthRatio`.

'''

from __future__ import annotations

import itertools

import numpy as np


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
