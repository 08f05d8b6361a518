'''Random particle clouds for the sum checks, after the reference's tests/testSphGpu.py:32-62 recipe.'''

import numpy as np

from oracle import array_formulation as af

rho0, c0, alpha = 1000.0, 20.0, 0.1


def cloud(dimensions: int, periodic: bool = False, seed: int = 3, kernelType: str = 'wendlandC2') -> dict:
    '''600 (2D) or 900 (3D) particles 25 m from the origin, 80% fluid, densities within 1%, masses within 5%.'''
    rng = np.random.default_rng(seed)
    count, spacing = (600, 0.04) if dimensions == 2 else (900, 0.045)
    extent = np.full(dimensions, spacing * count**(1.0 / dimensions))
    positions = 25.0 + rng.uniform(0.0, 1.0, size = (count, dimensions)) * extent
    period = None
    if periodic:
        positions[:, 1] -= 25.0
        period = float(extent[1])
    h = (2.0 if kernelType == 'wendlandC2' else 1.3) * spacing
    velocities = rng.normal(0.0, 1.0, size = (count, dimensions))
    densities = rho0 * (1.0 + rng.uniform(-0.01, 0.01, size = count))
    masses = rho0 * spacing**dimensions * (1.0 + rng.uniform(-0.05, 0.05, size = count))
    isFluid = rng.uniform(size = count) < 0.8
    pressures = af.taitEquationOfState(densities, rho0, c0)
    i, j = af.neighborPairs(positions, 2.0 * h, period)
    separation = positions[i] - positions[j]
    if period is not None:
        separation[:, 1] -= period * np.round(separation[:, 1] / period)
    distance = np.linalg.norm(separation, axis = 1)
    return dict(kernelType = kernelType, positions = positions, period = period, h = h, velocities = velocities,
                densities = densities, masses = masses, isFluid = isFluid, pressures = pressures, i = i, j = j,
                separation = separation, distance = distance, dims = dimensions,
                values = af.kernelValues(kernelType, distance, h, dimensions),
                gradients = af.kernelGradients(kernelType, separation, distance, h, dimensions))


def strainedFlow(case: dict) -> np.ndarray:
    '''A linear flow under noise over the cloud (tests/testSphGpu.py:209-216).'''
    positions = case['positions']
    dimensions = positions.shape[1]
    return (positions - positions.mean(axis = 0)) @ np.random.default_rng(11).normal(size = (dimensions, dimensions)).T \
        + 0.3 * case['velocities']


def physicsFor(case: dict, **options):
    '''The oracle's Physics for a cloud, its grid sized from the cloud's extent as the reference's tests do.'''
    from oracle.oracle import Physics
    positions = case['positions']
    gravity = np.zeros(case['dims'])
    gravity[-1] = -9.81
    settings = dict(dims = case['dims'], h = case['h'], domainMin = positions.min(axis = 0),
                    domainMax = positions.max(axis = 0), kernelType = case['kernelType'],
                    period = case['period'] or 0.0, rho0 = rho0, c0 = c0, alpha = alpha, gravity = gravity)
    settings.update(options)
    return Physics(**settings)


def shiftingBlock(dimensions: int, kernelType: str) -> dict:
    '''

    A jittered block of water under a free surface 25 m from the origin, beside a wall of boundary
    particles, after the reference's tests/testSphGpu.py:321-343 recipe: interior, surface and
    next-to-surface particles all shift.

    '''
    rng = np.random.default_rng(4)
    spacing = 0.04
    h = (2.0 if kernelType == 'wendlandC2' else 1.3) * spacing
    upper = ([1.2, 0.64] if kernelType == 'wendlandC2' else [0.8, 0.4]) if dimensions == 2 else [0.56, 0.56, 0.4]
    axes = [np.arange(0.5 * spacing, u, spacing) for u in upper]
    grids = np.meshgrid(*axes, indexing = 'ij')
    positions = 25.0 + np.column_stack([g.ravel() for g in grids])
    positions += rng.uniform(-0.08, 0.08, size = positions.shape) * spacing
    isFluid = positions[:, 0] > 25.0 + 2.0 * spacing
    densities = rho0 * (1.0 + rng.uniform(-0.002, 0.002, size = len(positions)))
    masses = np.full(len(positions), rho0 * spacing**dimensions)
    i, j = af.neighborPairs(positions, 2.0 * h)
    separation = positions[i] - positions[j]
    distance = np.linalg.norm(separation, axis = 1)
    velocities = rng.normal(0.0, 0.3, size = positions.shape)
    timeStep, mach = 2e-4, 0.1
    return dict(kernelType = kernelType, positions = positions, period = None, h = h, spacing = spacing,
                velocities = velocities, densities = densities, masses = masses, isFluid = isFluid, i = i, j = j,
                separation = separation, distance = distance, dims = dimensions, timeStep = timeStep, mach = mach,
                size = (c0 * timeStep / h) * mach * (2.0 * h)**2,
                values = af.kernelValues(kernelType, distance, h, dimensions),
                gradients = af.kernelGradients(kernelType, separation, distance, h, dimensions),
                spacingKernel = float(af.kernelValues(kernelType, np.array([spacing]), h, dimensions)[0]))


def expectedShifts(block: dict) -> tuple:
    '''The array formulation's shifts, density changes, velocity changes and surface of a block.'''
    i, j, fluid = block['i'], block['j'], block['isFluid']
    fields = af.shiftingFields(block['masses'], block['densities'], i, j, block['separation'], block['values'],
                               block['gradients'], len(fluid), block['spacingKernel'])
    shift, change = af.shiftingDisplacement(*fields, i, j, block['values'], fluid, block['size'], block['h'],
                                            block['kernelType'])
    velocity = af.shiftedVelocityChange(block['masses'], block['densities'], block['velocities'], i, j,
                                        block['gradients'], shift, fluid[i] & fluid[j])
    surface = af.freeSurface(fields[1], fields[2], fluid, block['h'], block['kernelType'])
    return shift, change, velocity, surface
