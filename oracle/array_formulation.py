'''

The array formulation of pySurf's SPH equations, restated in numpy over explicit pair lists.

TEST INFRASTRUCTURE, NOT PRODUCT CODE: only tests/, `__graft_entry__.smoke()` and bench.py's
cpu_baseline leg may import anything under oracle/, and only as the checker.

Each function restates one function of the reference's numpy backend (sphSolver.py), the float64
form of the step. Every pair appears once as (i, j) with i < j and is scattered to both
members, the structure of sphSolver.py, which differs from the per-particle gathers of the C++
restatement (oracle/sph_oracle.cpp) and of the device. Comparing the two therefore checks index,
sign and symmetry mistakes that a retyped formula would repeat.

'''

from __future__ import annotations

import math

import numpy as np
from scipy.spatial import cKDTree

taitExponent = 7.0           # sphSolver.py:225
viscositySoftening = 0.01    # sphSolver.py:229
limiterCloseDistance = 0.3   # sphSolver.py:234
limiterFoldWidth = 0.2       # sphSolver.py:235
correctionFloor = 0.25       # sphSolver.py:262


def sigmaUnit(kernelType: str, dimensions: int) -> float:
    '''Kernel normalization at h = 1 (sphKernels.py:439-442, sphGpu.py:1424-1444).'''
    if kernelType == 'wendlandC2':
        return 7.0 / (4.0 * math.pi) if dimensions == 2 else 21.0 / (16.0 * math.pi)
    return 10.0 / (7.0 * math.pi) if dimensions == 2 else 1.0 / math.pi


def kernelValues(kernelType: str, r: np.ndarray, h, dimensions: int) -> np.ndarray:
    '''W(r, h), h one value or one per pair (sphKernels.py:387-644).'''
    r = np.asarray(r, dtype = float)
    h = np.broadcast_to(np.asarray(h, dtype = float), r.shape)
    q = r / h
    sigma = sigmaUnit(kernelType, dimensions) / h**dimensions
    if kernelType == 'wendlandC2':
        rest = np.maximum(1.0 - 0.5 * q, 0.0)
        return np.where(q < 2.0, sigma * rest**4 * (2.0 * q + 1.0), 0.0)
    inner = sigma * (1.0 - 1.5 * q**2 + 0.75 * q**3)
    outer = sigma * 0.25 * np.maximum(2.0 - q, 0.0)**3
    return np.where(q < 1.0, inner, np.where(q < 2.0, outer, 0.0))


def kernelSlopes(kernelType: str, r: np.ndarray, h, dimensions: int) -> np.ndarray:
    '''dW/dr (sphKernels.py:387-644).'''
    r = np.asarray(r, dtype = float)
    h = np.broadcast_to(np.asarray(h, dtype = float), r.shape)
    q = r / h
    sigma = sigmaUnit(kernelType, dimensions) / h**dimensions
    if kernelType == 'wendlandC2':
        rest = np.maximum(1.0 - 0.5 * q, 0.0)
        return np.where(q < 2.0, -5.0 * sigma * q * rest**3 / h, 0.0)
    inner = sigma * (-3.0 * q + 2.25 * q**2) / h
    outer = sigma * (-0.75 * np.maximum(2.0 - q, 0.0)**2) / h
    return np.where(q < 1.0, inner, np.where(q < 2.0, outer, 0.0))


def kernelGradients(kernelType: str, separation: np.ndarray, distance: np.ndarray, h, dimensions: int) -> np.ndarray:
    '''grad_i W_ij = dW/dr (r_i - r_j) / r (sphKernels.py:607-644).'''
    slope = kernelSlopes(kernelType, distance, h, dimensions)
    safe = np.where(distance > 1e-12, distance, 1.0)
    return (slope / safe)[:, None] * separation


def neighborPairs(positions: np.ndarray, radius: float, period: float | None = None) -> tuple[np.ndarray, np.ndarray]:
    '''Every pair i < j closer than `radius`, y periodic when a period is given (sphNeighbors.py:28-78).'''
    if len(positions) < 2:
        empty = np.array([], dtype = np.intp)
        return empty, empty.copy()
    if period is None:
        tree = cKDTree(positions)
    else:
        data = positions - positions.min(axis = 0)
        box = 2.0 * data.max(axis = 0) + 4.0 * radius + 1.0
        wrapped = np.mod(positions[:, 1], period)
        wrapped[wrapped >= period] = 0.0
        data[:, 1] = wrapped
        box[1] = period
        tree = cKDTree(data, boxsize = box)
    pairs = tree.query_pairs(radius, output_type = 'ndarray')
    if pairs.size == 0:
        empty = np.array([], dtype = np.intp)
        return empty, empty.copy()
    return pairs[:, 0].astype(np.intp), pairs[:, 1].astype(np.intp)


def bruteForcePairs(positions: np.ndarray, lengths: np.ndarray, period: float | None = None) -> set:
    '''Every ordered pair (i, j), j != i, with r_ij < h_i + h_j, by brute force in float64.'''
    n = len(positions)
    result = set()
    for i in range(n):
        d = positions[i] - positions
        if period is not None:
            d[:, 1] -= period * np.round(d[:, 1] / period)
        r = np.sqrt(np.sum(d * d, axis = 1))
        hits = np.flatnonzero(r < lengths[i] + lengths)
        result.update((i, int(j)) for j in hits if j != i)
    return result


def scatterAdd(target: np.ndarray, index: np.ndarray, values: np.ndarray) -> np.ndarray:
    '''target[index[k]] += values[k] (sphSolver.py:573-592).'''
    if np.ndim(values) == 1:
        np.add.at(target, index, values)
    else:
        for column in range(values.shape[1]):
            target[:, column] += np.bincount(index, weights = values[:, column], minlength = target.shape[0])
    return target


def taitEquationOfState(density, referenceDensity: float, soundSpeed: float, clampNegative: bool = True):
    '''p = B ((rho / rho0)^7 - 1), B = rho0 c0^2 / 7 (sphSolver.py:541-557).'''
    stiffness = referenceDensity * soundSpeed**2 / taitExponent
    pressure = stiffness * ((np.asarray(density) / referenceDensity)**taitExponent - 1.0)
    return np.maximum(pressure, 0.0) if clampNegative else pressure


def taitDensity(pressure, referenceDensity: float, soundSpeed: float):
    '''The inverse of the Tait equation (sphSolver.py:559-566).'''
    stiffness = referenceDensity * soundSpeed**2 / taitExponent
    return referenceDensity * (1.0 + np.asarray(pressure) / stiffness)**(1.0 / taitExponent)


def continuityRate(masses, velocities, i, j, gradients, count, active = None, corrections = None):
    '''sum_j m_j (v_i - v_j) . grad W_ij, optionally with L_i (sphSolver.py:623-654).'''
    difference = velocities[i] - velocities[j]
    if corrections is None:
        rateI = rateJ = np.sum(difference * gradients, axis = 1)
    else:
        rateI = np.einsum('pa,pab,pb->p', difference, corrections[i], gradients)
        rateJ = np.einsum('pa,pab,pb->p', difference, corrections[j], gradients)
    if active is not None:
        rateI, rateJ = np.where(active, rateI, 0.0), np.where(active, rateJ, 0.0)
    rate = np.zeros(count)
    scatterAdd(rate, i, masses[j] * rateI)
    scatterAdd(rate, j, masses[i] * rateJ)
    return rate


def correctionMatrices(masses, densities, i, j, separation, gradients, isFluid):
    '''L_i = M_i^-1, M_i = sum_j V_j (x_j - x_i) (x) grad_i W_ij (sphSolver.py:656-689).'''
    count, dimensions = len(densities), separation.shape[1]
    volumes = masses / densities
    outer = -separation[:, :, None] * gradients[:, None, :]
    matrices = np.zeros((count, dimensions, dimensions))
    for a in range(dimensions):
        for b in range(dimensions):
            column = np.zeros(count)
            scatterAdd(column, i, volumes[j] * outer[:, a, b])
            scatterAdd(column, j, volumes[i] * outer[:, a, b])
            matrices[:, a, b] = column
    corrections = np.broadcast_to(np.eye(dimensions), matrices.shape).copy()
    applies = isFluid & (np.linalg.det(matrices) >= correctionFloor**dimensions)
    corrections[applies] = np.linalg.inv(matrices[applies])
    return corrections


def renormalizedDensityGradient(masses, densities, i, j, gradients, corrections, isFluid):
    '''L_i sum_j V_j (rho_j - rho_i) grad_i W_ij over every neighbor, zero on the boundary (sphSolver.py:691-716).'''
    count = len(densities)
    volumes = np.broadcast_to(np.asarray(masses, dtype = float), (count,)) / densities
    difference = densities[j] - densities[i]
    result = np.zeros((count, gradients.shape[1]))
    # grad_j W_ji = -grad_i W_ij and rho_i - rho_j = -difference: both members take the same sign
    scatterAdd(result, i, (volumes[j] * difference)[:, None] * np.einsum('pab,pb->pa', corrections[i], gradients))
    scatterAdd(result, j, (volumes[i] * difference)[:, None] * np.einsum('pab,pb->pa', corrections[j], gradients))
    result[~np.asarray(isFluid, dtype = bool)] = 0.0
    return result


def hydrostaticDensityDifference(vertical, meanDensity, referenceDensity, soundSpeed, gravity):
    '''rho0 (g / c0^2) (z_i - z_j) (rho_bar / rho0)^(2 - gamma) (sphSolver.py:718-737).'''
    ratio = np.asarray(meanDensity) / referenceDensity
    return referenceDensity * gravity / soundSpeed**2 * np.asarray(vertical) * ratio**(2.0 - taitExponent)


def densityDiffusionRate(masses, densities, i, j, separation, distance, gradients, count, h, soundSpeed, delta,
                         referenceDensity, gravity = None, active = None, densityGradients = None):
    '''

    The delta-SPH diffusive term, Molteni's, with the hydrostatic part removed, or with Antuono's
    1/2 (grad rho_i + grad rho_j) . (r_j - r_i) removed given the renormalized `densityGradients`
    (sphSolver.py:737-822).

    '''
    if active is not None:
        i, j = i[active], j[active]
        separation, distance, gradients = separation[active], distance[active], gradients[active]
        if np.ndim(h):
            h = h[active]
        if np.ndim(delta):
            delta = delta[active]
    difference = densities[j] - densities[i]
    if gravity is not None:
        difference = difference - hydrostaticDensityDifference(separation[:, -1], 0.5 * (densities[i] + densities[j]),
                                                               referenceDensity, soundSpeed, gravity)
    elif densityGradients is not None:
        difference = difference + 0.5 * np.sum((densityGradients[i] + densityGradients[j]) * separation, axis = 1)
    outward = -np.sum(separation * gradients, axis = 1)
    weight = 2.0 * delta * h * soundSpeed * difference * outward / (distance**2 + viscositySoftening * h**2)
    volumes = masses / densities
    rate = np.zeros(count)
    scatterAdd(rate, i, volumes[j] * weight)
    scatterAdd(rate, j, -volumes[i] * weight)
    return rate


def extrapolatedWallPressure(densities, pressures, i, j, separation, values, isFluid, gravity):
    '''Adami et al.'s wall pressure on every boundary particle (sphSolver.py:1045-1106).'''
    count = len(pressures)
    wallFirst = ~isFluid[i] & isFluid[j]
    fluidFirst = isFluid[i] & ~isFluid[j]
    wall = np.concatenate([i[wallFirst], j[fluidFirst]])
    water = np.concatenate([j[wallFirst], i[fluidFirst]])
    toWall = np.concatenate([separation[wallFirst], -separation[fluidFirst]])
    weights = np.concatenate([values[wallFirst], values[fluidFirst]])
    weight = np.zeros(count)
    total = np.zeros(count)
    scatterAdd(weight, wall, weights)
    scatterAdd(total, wall, weights * (pressures[water] + densities[water] * (toWall @ np.asarray(gravity))))
    result = pressures.copy()
    boundary = ~isFluid
    reached = boundary & (weight > 0.0)
    result[boundary] = 0.0
    result[reached] = np.maximum(total[reached] / weight[reached], 0.0)
    return result


def shepardDensity(masses, densities, i, j, values, selfValues):
    '''sum_j m_j W_ij / sum_j (m_j / rho_j) W_ij, each particle's own W(0, h_i) included (sphSolver.py:1023-1043).'''
    count = len(densities)
    masses = np.broadcast_to(np.asarray(masses, dtype = float), (count,))
    volumes = masses / densities
    numerator = masses * selfValues
    denominator = volumes * selfValues
    scatterAdd(numerator, i, masses[j] * values)
    scatterAdd(numerator, j, masses[i] * values)
    scatterAdd(denominator, i, volumes[j] * values)
    scatterAdd(denominator, j, volumes[i] * values)
    return numerator / denominator


def xsphVelocity(velocities, masses, densities, i, j, values, epsilon):
    '''v_i + epsilon sum_j (m_j / rho_bar_ij) (v_j - v_i) W_ij over the pairs given (sphSolver.py:1358-1377).'''
    count = len(densities)
    masses = np.broadcast_to(np.asarray(masses, dtype = float), (count,))
    weight = values / (0.5 * (densities[i] + densities[j]))
    difference = velocities[j] - velocities[i]
    correction = np.zeros_like(velocities)
    scatterAdd(correction, i, (masses[j] * weight)[:, None] * difference)
    scatterAdd(correction, j, -(masses[i] * weight)[:, None] * difference)
    return velocities + epsilon * correction


closureSmagorinsky = 0.12    # sphSolver.py:279
closureDiffusion = 1.5       # sphSolver.py:280
closureFilterRatio = 2.0     # sphSolver.py:281
closureCap = 0.2             # sphSolver.py:282


antiClumpingWeight = 0.2                                      # sphSolver.py:258
freeSurfaceDivergence = {2: 1.5, 3: 2.4}                      # sphSolver.py:265
freeSurfaceLopsidedness = {'cubicSpline': 0.3, 'wendlandC2': 0.4}  # sphSolver.py:266
surfaceGradient = 0.03                                        # sphSolver.py:271


def shiftingFields(masses, densities, i, j, separation, values, gradients, count, spacingKernel):
    '''grad C with the anti-clumping weight, grad C0, div r and grad rho over the pairs (sphSolver.py:824-887).'''
    volumes = np.broadcast_to(np.asarray(masses, dtype = float), (count,)) / densities
    weight = 1.0 + antiClumpingWeight * (values / spacingKernel)**4
    dimensions = separation.shape[1]
    gradient, plain = np.zeros((count, dimensions)), np.zeros((count, dimensions))
    scatterAdd(gradient, i, (volumes[j] * weight)[:, None] * gradients)
    scatterAdd(gradient, j, -(volumes[i] * weight)[:, None] * gradients)
    scatterAdd(plain, i, volumes[j][:, None] * gradients)
    scatterAdd(plain, j, -volumes[i][:, None] * gradients)
    term = -np.sum(separation * gradients, axis = 1)
    divergence = np.zeros(count)
    scatterAdd(divergence, i, volumes[j] * term)
    scatterAdd(divergence, j, volumes[i] * term)
    difference = (densities[j] - densities[i])[:, None] * gradients
    densityGradient = np.zeros((count, dimensions))
    scatterAdd(densityGradient, i, volumes[j][:, None] * difference)
    scatterAdd(densityGradient, j, volumes[i][:, None] * difference)
    return gradient, plain, divergence, densityGradient


def freeSurface(plain, divergence, isFluid, h, kernelType):
    '''div r below 1.5 (2D) or 2.4 (3D), or h |grad C0| above the kernel's lopsidedness (sphSolver.py:889-902).'''
    lopsided = h * np.linalg.norm(plain, axis = 1) > freeSurfaceLopsidedness[kernelType]
    return np.asarray(isFluid, dtype = bool) & ((divergence < freeSurfaceDivergence[plain.shape[1]]) | lopsided)


def shiftingDisplacement(gradient, plain, divergence, densityGradient, i, j, values, isFluid, size, h, kernelType):
    '''The shift, zero on the surface, projected beside it, and its density change (sphSolver.py:904-959).'''
    fluid = np.asarray(isFluid, dtype = bool)
    surface = freeSurface(plain, divergence, fluid, h, kernelType)
    length = np.linalg.norm(plain, axis = 1)
    surfaceNormal = np.zeros_like(plain)
    surfaceNormal[surface] = -plain[surface] / length[surface, None]
    normal = np.zeros_like(plain)
    scatterAdd(normal, i, values[:, None] * surfaceNormal[j])
    scatterAdd(normal, j, values[:, None] * surfaceNormal[i])
    vicinity = np.zeros(len(fluid), dtype = bool)
    vicinity[i[surface[j]]] = True
    vicinity[j[surface[i]]] = True
    weight = np.linalg.norm(normal, axis = 1)
    normal = np.divide(normal, weight[:, None], out = np.zeros_like(normal), where = weight[:, None] > 0.0)
    shift = -np.reshape(size, (-1, 1)) * gradient
    shift[vicinity] -= np.sum(shift[vicinity] * normal[vicinity], axis = 1)[:, None] * normal[vicinity]
    strong = vicinity & (h * np.linalg.norm(gradient, axis = 1) > surfaceGradient)
    shift[surface | strong | ~fluid] = 0.0
    return shift, np.sum(shift * densityGradient, axis = 1)


def shiftedVelocityChange(masses, densities, velocities, i, j, gradients, shift, active):
    '''sum_j V_j (v_j - v_i) (grad W_ij . delta r_i) over the active pairs (sphSolver.py:961-1010).'''
    count = len(densities)
    volumes = np.broadcast_to(np.asarray(masses, dtype = float), (count,)) / densities
    i, j, gradients = i[active], j[active], gradients[active]
    difference = velocities[j] - velocities[i]
    change = np.zeros_like(velocities)
    scatterAdd(change, i, (volumes[j] * np.sum(gradients * shift[i], axis = 1))[:, None] * difference)
    scatterAdd(change, j, (volumes[i] * np.sum(gradients * shift[j], axis = 1))[:, None] * difference)
    return change


def harmonicMean(a, b):
    '''2 a b / (a + b), zero where both are zero.'''
    both = a + b
    return np.where(both > 0.0, 2.0 * a * b / np.where(both > 0.0, both, 1.0), 0.0)


def closureCoefficients(gradients, isFluid, h, soundSpeed):
    '''The delta-LES-SPH alpha_i = K (C_S l)^2 ||D|| / (c0 h), delta_i = (C_delta l)^2 ||D|| / (c0 h), capped (sphSolver.py:1240-1281).'''
    dimensions = gradients.shape[1]
    strain = 0.5 * (gradients + np.swapaxes(gradients, 1, 2))
    norm = np.sqrt(2.0 * np.sum(strain * strain, axis = (1, 2)))
    length = closureFilterRatio * np.broadcast_to(np.asarray(h, dtype = float), norm.shape)
    scale = length**2 * norm / (soundSpeed * np.asarray(h, dtype = float))
    alphas = np.minimum(2.0 * (dimensions + 2) * closureSmagorinsky**2 * scale, closureCap)
    deltas = np.minimum(closureDiffusion**2 * scale, closureCap)
    fluid = np.asarray(isFluid, dtype = bool)
    return np.where(fluid, alphas, 0.0), np.where(fluid, deltas, 0.0)


def closureViscosity(relative, separation, distance, densitiesI, densitiesJ, h, soundSpeed, referenceDensity, alphasI,
                     alphasJ, kinematicViscosity):
    '''Pi_ij = -alpha_ij c0 h rho0 (v_ij . r_ij) / (r^2 rho_i rho_j), alpha_ij = K nu / (c0 h) + mean (sphSolver.py:1283-1330).'''
    dimensions = separation.shape[1]
    pairAlpha = 2.0 * (dimensions + 2) * kinematicViscosity / (soundSpeed * h) + harmonicMean(alphasI, alphasJ)
    approach = np.sum(relative * separation, axis = 1)
    return -pairAlpha * soundSpeed * h * referenceDensity * approach / (distance**2 * densitiesI * densitiesJ)


def artificialViscosity(relative, separation, distance, meanDensity, h, soundSpeed, alpha, reconstructed = None):
    '''Monaghan's Pi_ij, acting on approaching pairs (sphSolver.py:1122-1149).'''
    approach = np.sum(relative * separation, axis = 1)
    acting = approach < 0.0
    if reconstructed is not None:
        approach = np.sum(reconstructed * separation, axis = 1)
        acting &= approach < 0.0
    mu = h * approach / (distance**2 + viscositySoftening * h**2)
    return np.where(acting, -alpha * soundSpeed * mu / meanDensity, 0.0)


def velocityGradients(masses, densities, velocities, i, j, separation, gradients, isFluid):
    '''G_i = S_i M_i^-1 over fluid pairs (sphSolver.py:1151-1190).'''
    count, dimensions = len(densities), separation.shape[1]
    volumes = masses / densities
    fluidPairs = isFluid[i] & isFluid[j]
    i, j = i[fluidPairs], j[fluidPairs]
    gradient = gradients[fluidPairs]
    rise = (velocities[j] - velocities[i])[:, :, None] * gradient[:, None, :]
    offset = -separation[fluidPairs][:, :, None] * gradient[:, None, :]
    sums = np.zeros((count, dimensions, dimensions))
    matrices = np.zeros((count, dimensions, dimensions))
    for a in range(dimensions):
        for b in range(dimensions):
            for target, outer in ((sums, rise), (matrices, offset)):
                column = np.zeros(count)
                scatterAdd(column, i, volumes[j] * outer[:, a, b])
                scatterAdd(column, j, volumes[i] * outer[:, a, b])
                target[:, a, b] = column
    result = np.zeros((count, dimensions, dimensions))
    applies = isFluid & (np.linalg.det(matrices) >= correctionFloor**dimensions)
    result[applies] = sums[applies] @ np.linalg.inv(matrices[applies])
    return result


def limitedVelocityDifference(relative, separation, distance, h, gradientsI, gradientsJ):
    '''v_ij less (phi_ij / 2) (G_i + G_j) r_ij under MAGMA2's limiter (sphSolver.py:1192-1230).'''
    rampI = np.einsum('pab,pb->pa', gradientsI, separation)
    rampJ = np.einsum('pab,pb->pa', gradientsJ, separation)
    slopeI = np.sum(separation * rampI, axis = 1)
    slopeJ = np.sum(separation * rampJ, axis = 1)
    agree = slopeI * slopeJ > 0.0
    ratio = np.divide(slopeI, slopeJ, out = np.zeros_like(slopeI), where = agree)
    limiter = np.where(agree, np.minimum(1.0, 4.0 * ratio / (1.0 + ratio)**2), 0.0)
    eta = distance / h
    limiter *= np.where(eta < limiterCloseDistance, np.exp(-((eta - limiterCloseDistance) / limiterFoldWidth)**2), 1.0)
    return relative - 0.5 * limiter[:, None] * (rampI + rampJ)


def pairAccelerations(masses, densities, pressures, i, j, gradients, viscosity, count):
    '''-sum_j m_j (p_i / rho_i^2 + p_j / rho_j^2 + Pi_ij) grad W_ij (sphSolver.py:1328-1356).'''
    coefficient = -(pressures[i] / densities[i]**2 + pressures[j] / densities[j]**2 + viscosity)
    accelerations = np.zeros((count, gradients.shape[1]))
    scatterAdd(accelerations, i, (masses[j] * coefficient)[:, None] * gradients)
    scatterAdd(accelerations, j, -(masses[i] * coefficient)[:, None] * gradients)
    return accelerations


def cellOrderKey(positions, cellWidth: float, shape) -> np.ndarray:
    '''Cells from the origin, wrapped onto the grid, x fastest (sphSolver.py:1379-1408).'''
    cells = np.floor(np.asarray(positions, dtype = float) / cellWidth).astype(np.int64)
    nx, ny, nz = (int(count) for count in shape)
    x, y = np.mod(cells[:, 0], nx), np.mod(cells[:, 1], ny)
    z = np.mod(cells[:, 2], nz) if cells.shape[1] == 3 else np.zeros_like(x)
    return (z * ny + y) * nx + x


def cflTimeStep(h, soundSpeed, maxSpeed, maxAcceleration, cflNumber, maxTimeStep) -> float:
    '''min(CFL h / (c0 + |v|max), CFL sqrt(h / |a|max), maxTimeStep) (sphSolver.py:1410-1424).'''
    timeStep = cflNumber * h / (soundSpeed + maxSpeed)
    if maxAcceleration > 1e-12:
        timeStep = min(timeStep, cflNumber * math.sqrt(h / maxAcceleration))
    return min(timeStep, maxTimeStep)


def gridDimensions(domainMin, domainMax, cellSize: float, dimensions: int, maxCells: int = 2**25) -> tuple:
    '''Cells along x, y, z covering the domain, one support across (sphGpu.py:1466-1478).'''
    extent = np.zeros(3)
    extent[:dimensions] = np.asarray(domainMax, dtype = float) - np.asarray(domainMin, dtype = float)
    cells = np.maximum(np.ceil(extent / cellSize).astype(int) + 3, 1)
    if dimensions == 2:
        cells[2] = 1
    while int(np.prod(cells)) > maxCells:
        cells = np.maximum((cells * 0.8).astype(int), 1)
    return tuple(int(count) for count in cells)


def elasticEnergy(densities, referenceDensity: float, soundSpeed: float, tension: bool = False) -> np.ndarray:

    '''

    The elastic energy per unit mass the Tait equation stores at a density, the integral of p / s^2
    from rho0 to rho (sphEnergy.py:120-160): (c0^2 / 7) x^2 (21 + 35x + 35x^2 + 21x^3 + 7x^4 + x^5)
    / (6 (1 + x)) with x = rho / rho0 - 1; zero below rho0 unless the fluid carries tension.

    '''

    x = np.asarray(densities, dtype = float) / referenceDensity - 1.0
    polynomial = 21.0 + x * (35.0 + x * (35.0 + x * (21.0 + x * (7.0 + x))))
    energy = soundSpeed**2 / 7.0 * x * x * polynomial / (6.0 * (1.0 + x))
    return energy if tension else np.where(x < 0.0, 0.0, energy)


def globalSums(positions, velocities, densities, masses, isFluid, gravity: float, referenceDensity: float,
               soundSpeed: float, tension: bool = False) -> dict:

    '''

    Float64 sums over the fluid: mass, kinetic energy sum m |v|^2 / 2 and potential energy sum m g z
    (sphParticles.py:236-255, z the last axis), elastic energy sum m e(rho) and momentum sum m v.

    '''

    fluid = np.asarray(isFluid, dtype = bool)
    m = np.asarray(masses, dtype = float)[fluid]
    v = np.asarray(velocities, dtype = float)[fluid]
    return {'mass': float(m.sum()), 'kinetic': float(0.5 * np.sum(m * np.sum(v * v, axis = 1))),
            'potential': float(np.sum(m * gravity * np.asarray(positions, dtype = float)[fluid, -1])),
            'elastic': float(np.sum(m * elasticEnergy(np.asarray(densities)[fluid], referenceDensity, soundSpeed,
                                                      tension))),
            'momentum': (m[:, None] * v).sum(axis = 0)}
