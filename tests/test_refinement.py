'''

Splitting and merging on the host (pysurf_amd/refinement.py) on the arrangements of
tests/testSphRefinement.py:30-215: a split conserves mass, momentum, kinetic energy and the centre
of mass and lays the daughters on the mother's lattice refined by two; a merge conserves mass,
momentum and the centre of mass, removes exactly the kinetic energy it reports, lands no particle
below its level, restores a mother split just before, keeps the boundary's indices, and leaves
particles that never split alone.

'''

import numpy as np
import pytest

from pysurf_amd.particles import ParticleSystem, latticePoints
from pysurf_amd.refinement import compactionSource, mergeParticles, refinementLevels, splitParticles

rho0, c0 = 1000.0, 20.0


def lattice(dimensions, spacing = 0.02, seed = 5):
    rng = np.random.default_rng(seed)
    upper = [0.3, 0.2] if dimensions == 2 else [0.2, 0.12, 0.12]
    fluid = latticePoints(np.zeros(dimensions), np.array(upper), spacing)
    floor = latticePoints(np.r_[np.zeros(dimensions - 1), -spacing], np.r_[upper[:-1], 0.0], spacing)
    particles = ParticleSystem.assemble(fluid, floor, spacing, rho0, rng.normal(0.0, 1.0, size = fluid.shape))
    particles.densities = rho0 * (1.0 + rng.uniform(-0.01, 0.01, size = particles.nParticles))
    particles.pressures = rho0 * c0**2 / 7.0 * ((particles.densities / rho0)**7 - 1.0)
    particles.accelerations = rng.normal(0.0, 1.0, size = particles.positions.shape)
    return particles


def totals(particles):
    fluid = particles.isFluid
    m = particles.masses[fluid]
    return (m.sum(), (m[:, None] * particles.velocities[fluid]).sum(axis = 0), particles.kineticEnergy(),
            (m[:, None] * particles.positions[fluid]).sum(axis = 0))


@pytest.mark.parametrize('dimensions', [2, 3])
def testASplitConservesMassMomentumEnergyAndTheCentreOfMass(dimensions):
    particles = lattice(dimensions)
    before = totals(particles)
    mask = particles.positions[:, 0] > 0.1
    chosen = int(np.sum(mask & particles.isFluid))
    count = particles.nParticles
    assert splitParticles(particles, mask, rho0, 0.05) == chosen > 0
    after = totals(particles)
    assert particles.nParticles == count + chosen * (2**dimensions - 1)
    assert after[0] == pytest.approx(before[0], rel = 1e-14)
    assert after[1] == pytest.approx(before[1], rel = 1e-12, abs = 1e-12)
    assert after[2] == pytest.approx(before[2], rel = 1e-13)
    assert after[3] == pytest.approx(before[3], rel = 1e-13)


@pytest.mark.parametrize('dimensions', [2, 3])
def testAMergeConservesMassMomentumAndTheCentreOfMassAndReportsTheEnergyItRemoves(dimensions):
    spacing = 0.02
    particles = lattice(dimensions, spacing)
    count = particles.nParticles
    splitParticles(particles, particles.positions[:, 0] > 0.1, rho0, 0.05)
    rng = np.random.default_rng(3)
    daughters = refinementLevels(particles, rho0 * spacing**dimensions) == 1
    particles.positions[daughters] += rng.uniform(-0.1, 0.1, size = (int(daughters.sum()), dimensions)) * spacing
    particles.velocities[daughters] += rng.normal(0.0, 0.3, size = (int(daughters.sum()), dimensions))
    before = totals(particles)
    split = particles.nParticles
    result = mergeParticles(particles, np.ones(particles.nParticles, dtype = bool), rho0, rho0 * spacing**dimensions, 0.05)
    after = totals(particles)
    assert result.groups > 0.8 * (split - count) / (2**dimensions - 1)
    assert particles.nParticles == split - result.groups * (2**dimensions - 1)
    assert after[0] == pytest.approx(before[0], rel = 1e-14)
    assert after[1] == pytest.approx(before[1], rel = 1e-12, abs = 1e-12)
    assert after[3] == pytest.approx(before[3], rel = 1e-13)
    assert result.removedEnergy > 0.0
    assert after[2] == pytest.approx(before[2] - result.removedEnergy, rel = 1e-12)
    assert set(np.unique(refinementLevels(particles, rho0 * spacing**dimensions)[particles.isFluid])) <= {0, 1}


@pytest.mark.parametrize('dimensions', [2, 3])
def testAMergeStraightAfterASplitRestoresTheMother(dimensions):
    spacing, length = 0.02, 0.06
    particles = lattice(dimensions, spacing)
    count = particles.nParticles
    mother = int(np.flatnonzero(particles.isFluid)[7])
    state = {name: getattr(particles, name)[mother].copy() for name in ('positions', 'velocities', 'accelerations')}
    state.update(densities = particles.densities[mother], pressures = particles.pressures[mother],
                 masses = particles.masses[mother])
    others = particles.positions.copy()
    mask = np.zeros(count, dtype = bool)
    mask[mother] = True
    splitParticles(particles, mask, rho0, length)
    result = mergeParticles(particles, np.ones(particles.nParticles, dtype = bool), rho0, rho0 * spacing**dimensions, length)
    assert result.groups == 1 and result.removedEnergy == pytest.approx(0.0, abs = 1e-15)
    assert particles.nParticles == count
    for name, value in state.items():
        assert getattr(particles, name)[mother] == pytest.approx(value, rel = 1e-13, abs = 1e-15)
    assert particles.lengths[mother] == pytest.approx(length)
    assert particles.positions == pytest.approx(others, rel = 1e-13, abs = 1e-15)


def testAMergeKeepsTheBoundarysIndices():
    # Fluid first, then the floor, then the daughters: the groups' other members among the fluid's
    # places are filled from the end, so the floor keeps its slots
    spacing = 0.02
    particles = lattice(2, spacing)
    walls = ~particles.isFluid
    wallPositions = particles.positions[walls].copy()
    splitParticles(particles, particles.positions[:, 0] < 0.1, rho0, 0.05)
    result = mergeParticles(particles, np.ones(particles.nParticles, dtype = bool), rho0, rho0 * spacing**2, 0.05)
    assert result.groups > 0
    assert np.array_equal(~particles.isFluid[:len(walls)], walls)
    assert np.array_equal(particles.positions[:len(walls)][walls], wallPositions)


def testParticlesThatHaveNotSplitDoNotMerge():
    particles = lattice(2)
    count, positions = particles.nParticles, particles.positions.copy()
    result = mergeParticles(particles, np.ones(count, dtype = bool), rho0, rho0 * 0.02**2, 0.05)
    assert result.groups == 0 and particles.nParticles == count
    assert np.array_equal(result.source, np.arange(count)) and np.array_equal(particles.positions, positions)


def testTheCompactionFillsHolesAmongTheBoundaryFromTheEnd():
    # Slots 0-3 fluid, 4-5 boundary, 6-9 fluid; 1 and 7 leave: 1 takes 9, the rest keep their order
    fluid = np.array([1, 1, 1, 1, 0, 0, 1, 1, 1, 1], dtype = bool)
    assert compactionSource(10, np.array([1, 7]), fluid).tolist() == [0, 9, 2, 3, 4, 5, 6, 8]
    with pytest.raises(RuntimeError):
        compactionSource(6, np.array([0, 1, 2]), np.array([1, 1, 1, 0, 1, 1], dtype = bool))
