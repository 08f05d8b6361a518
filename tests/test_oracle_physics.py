'''

The reference's own physics checks of the pair sums and the step, applied to the float64
restatement (oracle/sph_oracle.cpp), and closed forms a plausible slip in the restatement fails:
a sign, a factor of two, a dropped term or a transposed operand.

Each test names the reference test it follows (tests/testSphSolver.py in the reference). They pin
the restatement directly, not through its agreement with the numpy restatement
(oracle/array_formulation.py), which is the same author's second reading of the same source.

  the viscosity limiter      leaves a linear flow alone, still damps noise, conserves momentum and
                             never feeds energy (:208-256); on a quadratic shear each pair's limiter
                             is min(1, 4A/(1+A)^2) of the two slopes' ratio A; its fade below
                             0.3 h is exp(-((q - 0.3)/0.2)^2) on a linear flow
  the viscosity Pi_ij        a closed form on two approaching particles; only approaching pairs
  the boundary factor        scales the wall pairs' viscosity only (:721-733)
  continuity                 -rho div v under an expansion (:282-298)
  the diffusive term         2 delta c0 h: delta c0 h lap(rho) on a quadratic density; a bump
                             flattens and mass is conserved, Molteni's and Antuono's
                             (:337-351, :463-483)
  the density floor          holds a droplet flying apart at the floor (:765-783)
  the periodic axis          keeps an extruded flow planar and its particles inside the period
                             (:793-816)
  particle shifting          leaves a lattice alone, parts a crowded pair, runs along the free
                             surface, carries the density and velocity of where it lands (:497-606)

'''

import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import oracle
from oracle.array_formulation import kernelGradients, kernelSlopes, kernelValues, sigmaUnit

rho0, c0 = 1000.0, 20.0
s = 0.01


def lattice(lower, upper, spacing: float = s) -> np.ndarray:
    '''Cell centres filling the box [lower, upper], half a spacing in from each face, x fastest.'''
    axes = [np.arange(low + spacing / 2.0, high, spacing) for low, high in zip(lower, upper)]
    grids = np.meshgrid(*axes, indexing = 'xy')
    return np.column_stack([grid.ravel() for grid in grids])


def physics(points: np.ndarray, h: float, kernelType: str = 'cubicSpline', **options) -> oracle.Physics:
    '''The restatement's physics for a cloud: its own grid around the points, no gravity unless given.'''
    dims = points.shape[1]
    settings = dict(dims = dims, h = h, domainMin = points.min(axis = 0) - 4.0 * h,
                    domainMax = points.max(axis = 0) + 4.0 * h, kernelType = kernelType, rho0 = rho0, c0 = c0,
                    gravity = np.zeros(dims))
    settings.update(options)
    return oracle.Physics(**settings)


def waterBlock(jitter: float = 0.0, seed: int = 4) -> np.ndarray:
    '''Water 40 spacings wide and 20 deep, free all round (the reference's block at h = 3 s, Wendland C2).'''
    rng = np.random.default_rng(seed)
    positions = lattice([0.0, 0.0], [40 * s, 20 * s])
    return positions + rng.uniform(-jitter * s, jitter * s, size = positions.shape)


def viscous(positions, velocities, limited: bool, alpha: float = 0.01, h: float = 3.0 * s,
            kernelType: str = 'wendlandC2', isFluid = None):
    '''Each particle's viscous acceleration at rest density (no pressure, no gravity), and the
    dissipation -sum m_i v_i . a_i it does [W].'''
    count = len(positions)
    fluid = np.ones(count, dtype = bool) if isFluid is None else isFluid
    masses, densities = np.full(count, rho0 * s * s), np.full(count, rho0)
    p = physics(positions, h, kernelType, alpha = alpha, limiter = limited)
    gradients = oracle.velocityGradients(p, positions, velocities, densities, masses, fluid) if limited else None
    acceleration = oracle.accelerations(p, positions, velocities, densities, np.zeros(count), masses, fluid,
                                        gradients = gradients)
    return acceleration, -np.sum(masses[:, None] * velocities * acceleration)


# -- The limited viscosity (tests/testSphSolver.py:208-256) -- #

def testTheLimitedViscosityLeavesALinearFlowAlone():
    positions = waterBlock()
    strain = positions @ np.array([[1.0, 0.0], [0.0, -1.0]]).T
    plain, plainPower = viscous(positions, strain, False)
    limited, limitedPower = viscous(positions, strain, True)
    assert np.count_nonzero(np.any(plain != 0.0, axis = 1)) > 0.4 * len(positions)
    assert np.abs(limited).max() < 1e-12 * np.abs(plain).max()
    assert limitedPower < 1e-12 * plainPower


def testTheLimitedViscosityStillDampsNoise():
    # noise of 0.2 spacings per second over a strain of one per second: the limited term dissipates at
    # least half what the plain term takes from the noise alone, and a tenth of what it takes from both
    positions = waterBlock()
    strain = positions @ np.array([[1.0, 0.0], [0.0, -1.0]]).T
    noise = np.random.default_rng(9).normal(scale = 0.2 * s, size = strain.shape)
    _, noisePower = viscous(positions, noise, False)
    _, plainPower = viscous(positions, strain + noise, False)
    _, limitedPower = viscous(positions, strain + noise, True)
    assert limitedPower > 0.5 * noisePower
    assert limitedPower < 0.1 * plainPower


def testTheLimitedViscosityConservesMomentumAndNeverFeedsEnergy():
    positions = waterBlock(jitter = 0.3, seed = 5)
    rng = np.random.default_rng(6)
    velocities = positions @ rng.normal(size = (2, 2)).T + rng.normal(scale = 0.05, size = positions.shape)
    acceleration, power = viscous(positions, velocities, True, alpha = 0.1)
    forces = rho0 * s * s * acceleration
    scale, count = np.abs(forces).max(), len(positions)
    assert scale > 0.0
    assert np.abs(forces.sum(axis = 0)).max() < 1e-12 * scale * count
    torque = np.sum(positions[:, 0] * forces[:, 1] - positions[:, 1] * forces[:, 0])
    assert abs(torque) < 1e-12 * scale * count
    assert power > 0.0


def piOfPair(approach: float, r: float, h: float, alpha: float, meanDensity: float) -> float:
    '''Monaghan's Pi_ij = -alpha c0 mu / rho_bar, mu = h (v_ij . r_ij) / (r^2 + 0.01 h^2), for an approaching pair.'''
    return -alpha * c0 * (h * approach / (r * r + 0.01 * h * h)) / meanDensity


def wendlandSlopeOverR(r: float, h: float) -> float:
    '''dW/dr / r of the 2D Wendland C2, sigma = 7 / (4 pi h^2): dW/dr = -5 sigma q (1 - q/2)^3 / h.'''
    q, sigma = r / h, 7.0 / (4.0 * math.pi * h * h)
    return -5.0 * sigma * q * (1.0 - 0.5 * q)**3 / h / r


def testTheViscosityOfTwoApproachingParticlesIsMonaghans():
    # a_1 = -m_2 Pi_12 grad_1 W(r) with Pi over the pair's mean density: the closed form, and nothing
    # when the pair separates (pressures held at zero, densities unequal)
    h, alpha = 1.3 * s, 0.1
    positions = np.array([[0.0, 0.0], [0.8 * s, 0.3 * s]])
    densities = np.array([1000.0, 1030.0])
    masses = np.full(2, rho0 * s * s)
    p = physics(positions, h, 'wendlandC2', alpha = alpha)
    d = positions[0] - positions[1]
    r = float(np.linalg.norm(d))
    for v1, approaching in (([0.5, 0.1], True), ([-0.5, -0.1], False)):
        velocities = np.array([v1, [0.0, 0.0]])
        acceleration = oracle.accelerations(p, positions, velocities, densities, np.zeros(2), masses, np.ones(2, dtype = bool))
        approach = float(np.dot(velocities[0] - velocities[1], d))
        assert (approach < 0.0) == approaching
        expected = -rho0 * s * s * piOfPair(approach, r, h, alpha, densities.mean()) * wendlandSlopeOverR(r, h) * d \
            if approaching else np.zeros(2)
        assert acceleration[0] == pytest.approx(expected, rel = 1e-12, abs = 1e-300)
        assert acceleration[1] == pytest.approx(-expected, rel = 1e-12, abs = 1e-300)
        if approaching:
            # the viscosity pushes the approaching particle back
            assert np.dot(acceleration[0], d) > 0.0


def testTheLimiterFadesBelowThreeTenthsOfH():
    # On a linear flow the limiter reconstructs every pair's velocity difference and Pi vanishes,
    # but on a pair closer than 0.3 h it fades by exp(-((q - 0.3) / 0.2)^2): of two particles 0.2 h
    # apart in a compressing flow, only that pair acts, with Pi of (1 - fade) of its approach
    h, alpha = 3.0 * s, 0.1
    positions = waterBlock()
    centre = int(np.argmin(np.linalg.norm(positions - [20 * s, 10 * s], axis = 1)))
    right = int(np.argmin(np.linalg.norm(positions - (positions[centre] + [s, 0.0]), axis = 1)))
    positions[centre, 0] += s - 0.2 * h
    field = np.array([[-0.8, 0.3], [0.5, 0.6]])
    velocities = positions @ field.T
    acceleration, _ = viscous(positions, velocities, True, alpha = alpha)
    d = positions[centre] - positions[right]
    r = float(np.linalg.norm(d))
    assert r / h == pytest.approx(0.2)
    fade = math.exp(-((r / h - 0.3) / 0.2)**2)
    approach = (1.0 - fade) * float(np.dot(field @ d, d))
    assert approach < 0.0
    expected = -rho0 * s * s * piOfPair(approach, r, h, alpha, rho0) * wendlandSlopeOverR(r, h) * d
    assert acceleration[centre] == pytest.approx(expected, rel = 1e-9)
    assert acceleration[right] == pytest.approx(-expected, rel = 1e-9)
    others = np.ones(len(positions), dtype = bool)
    others[[centre, right]] = False
    assert np.abs(acceleration[others]).max() < 1e-9 * np.abs(expected).max()


def testTheLimiterOfAQuadraticShearIsItsSlopeRatioFormula():
    # u = b x^2 on a lattice: G = [[2 b x, 0], [0, 0]] exactly away from the surface, so a pair's two
    # slopes are 2 b x_i dx^2 and 2 b x_j dx^2, A = x_i / x_j, and its limited approach is
    # b (x_i + x_j) dx^2 (1 - min(1, 4A / (1 + A)^2)): each interior particle's acceleration is a sum
    # of those closed forms over its neighbors
    # (a pair's two G are exact where both supports are full: 4 h inside the block)
    h, alpha, b = 3.0 * s, 0.1, -40.0
    positions = lattice([0.2, 0.0], [0.2 + 40 * s, 40 * s])
    velocities = np.column_stack([b * positions[:, 0]**2, np.zeros(len(positions))])
    acceleration, _ = viscous(positions, velocities, True, alpha = alpha)
    tree = cKDTree(positions)
    inside = positions - [0.2, 0.0]
    interior = np.flatnonzero(np.all((inside > 4 * h) & (inside < 40 * s - 4 * h), axis = 1))
    assert len(interior) > 50
    for i in interior[::7]:
        expected = np.zeros(2)
        for j in tree.query_ball_point(positions[i], 2.0 * h):
            d = positions[i] - positions[j]
            r = float(np.linalg.norm(d))
            if j == i or r >= 2.0 * h:
                continue
            xi, xj = positions[i, 0], positions[j, 0]
            ratio = xi / xj
            limiter = min(1.0, 4.0 * ratio / (1.0 + ratio)**2)
            approach = b * (xi + xj) * d[0]**2 * (1.0 - limiter)
            if approach < 0.0:
                expected += -rho0 * s * s * piOfPair(approach, r, h, alpha, rho0) * wendlandSlopeOverR(r, h) * d
        assert np.abs(expected).max() > 0.0
        assert acceleration[i] == pytest.approx(expected, rel = 1e-7, abs = 1e-9 * np.abs(expected).max())


# -- The boundary viscosity factor (tests/testSphSolver.py:721-733) -- #

def testTheBoundaryViscosityFactorScalesOnlyWallPairs():
    def first(otherIsFluid: bool, factor: float) -> np.ndarray:
        positions = np.array([[0.0, s], [0.0, 0.0]])
        velocities = np.array([[0.3, -0.5], [0.0, 0.0]])
        fluid = np.array([True, otherIsFluid])
        p = physics(positions, 1.3 * s, 'cubicSpline', alpha = 0.1, boundaryFactor = factor)
        a = oracle.accelerations(p, positions, velocities, np.full(2, rho0), np.zeros(2), np.full(2, rho0 * s * s), fluid)
        return a[0]

    assert first(False, 0.0) == pytest.approx([0.0, 0.0], abs = 1e-12)
    assert first(False, 1.0)[1] > 0.0
    assert first(False, 0.5) == pytest.approx(0.5 * first(False, 1.0), rel = 1e-14)
    assert first(True, 0.0) == pytest.approx(first(True, 1.0), rel = 1e-14)
    assert first(True, 0.0)[1] > 0.0
    # a fluid pair and a wall pair at factor one are the same pair
    assert first(True, 0.0) == pytest.approx(first(False, 1.0), rel = 1e-14)


# -- Continuity (tests/testSphSolver.py:282-298) -- #

def testContinuityGivesMinusRhoDivergenceForAnExpansion():
    # v = a x in 2D has div v = 2a, so d rho / dt = -2 a rho0 inside, to the cubic spline's lattice
    # error at h = 1.3 s (1.05 percent); a contraction gives the opposite sign
    positions = lattice([0.0, 0.0], [0.2, 0.2])
    count = len(positions)
    p = physics(positions, 1.3 * s)
    centre = int(np.argmin(np.linalg.norm(positions - 0.1, axis = 1)))
    for rate in (0.5, -0.3):
        result = oracle.continuityRate(p, positions, rate * positions, np.full(count, rho0), np.full(count, rho0 * s * s),
                                       np.ones(count, dtype = bool))
        assert result[centre] == pytest.approx(-2.0 * rate * rho0, rel = 0.02)


# -- The diffusive term (tests/testSphSolver.py:326-351, 463-483) -- #

def diffusionRate(positions, densities, mode: str = 'molteni', delta: float = 0.1, h: float = 1.3 * s,
                  kernelType: str = 'cubicSpline', corrections: bool = False):
    '''The diffusive term alone (fluid at rest, no gravity) of every particle.'''
    count = len(positions)
    fluid = np.ones(count, dtype = bool)
    masses = np.full(count, rho0 * s * s)
    p = physics(positions, h, kernelType, diffusion = mode, delta = delta)
    gradients = None
    if mode == 'antuono':
        L = oracle.correctionMatrices(p, positions, densities, masses, fluid)
        gradients = oracle.densityGradients(p, positions, densities, masses, fluid, L)
    return oracle.continuityRate(p, positions, np.zeros((count, 2)), densities, masses, fluid, densityGradients = gradients)


@pytest.mark.parametrize('kernelType, ratio', [('cubicSpline', 1.3), ('cubicSpline', 2.0), ('wendlandC2', 2.0),
                                                ('wendlandC2', 3.0)])
def testTheDiffusiveTermIsDeltaC0HTimesTheLaplacian(kernelType, ratio):
    # Molteni's term 2 delta c0 h sum_j V_j (rho_j - rho_i)(-r dW/dr) / (r^2 + 0.01 h^2) is
    # delta c0 h lap(rho) in the continuum (sum_j V_j (f_j - f_i)(-r W') / r^2 is half the Laplacian).
    # On rho0 + c x^2, lap(rho) = 2c: the interior rate is 2 delta c0 h c to the softening and the
    # lattice sum (within 4 percent), and twice delta doubles it exactly
    h, c = ratio * s, 2000.0
    positions = lattice([0.0, 0.0], [0.3, 0.3])
    densities = rho0 + c * (positions[:, 0] - 0.15)**2
    centre = int(np.argmin(np.linalg.norm(positions - 0.15, axis = 1)))
    rate = diffusionRate(positions, densities, h = h, kernelType = kernelType)
    assert rate[centre] == pytest.approx(2.0 * 0.1 * c0 * h * c, rel = 0.04)
    assert diffusionRate(positions, densities, delta = 0.2, h = h, kernelType = kernelType)[centre] == \
        pytest.approx(2.0 * rate[centre], rel = 1e-13)


def testAntuonosTermLeavesAnyLinearDensityAlone():
    # a density rising linearly along x is not water at rest: Molteni's form diffuses it where the
    # support is cut, Antuono's, which subtracts the renormalized gradients' projection, leaves it
    # alone at every particle to rounding (tests/testSphSolver.py:434-449)
    rng = np.random.default_rng(2)
    positions = lattice([0.0, 0.0], [0.2, 0.1]) + rng.uniform(-0.15, 0.15, size = (200, 2)) * s
    densities = rho0 + 50.0 * positions[:, 0]
    molteni = diffusionRate(positions, densities, 'molteni')
    antuono = diffusionRate(positions, densities, 'antuono')
    assert np.abs(molteni).max() > 1e-3 * rho0
    assert np.abs(antuono).max() < 1e-9 * np.abs(molteni).max()


@pytest.mark.parametrize('mode', ['molteni', 'antuono'])
def testDiffusionFlattensADensityBumpAndConservesMass(mode):
    # each pair moves its two densities toward each other: the bump falls, its neighbors rise, and
    # sum_i V_i D_i vanishes pair by pair (on jittered particles with random densities, too)
    positions = lattice([0.0, 0.0], [0.1, 0.1])
    centre = int(np.argmin(np.linalg.norm(positions - 0.05, axis = 1)))
    densities = np.full(len(positions), rho0)
    densities[centre] += 1.0
    rate = diffusionRate(positions, densities, mode)
    # Molteni's term raises every neighbor of the bump; Antuono's, whose gradients see the bump too,
    # its four nearest (tests/testSphSolver.py:478-483)
    reach = 2.0 * 1.3 * s if mode == 'molteni' else 1.01 * s
    near = (np.linalg.norm(positions - positions[centre], axis = 1) < reach) & (np.arange(len(positions)) != centre)
    assert near.sum() >= 4
    assert rate[centre] < 0.0
    assert np.all(rate[near] > 0.0)
    rng = np.random.default_rng(9)
    jittered = lattice([0.0, 0.0], [0.2, 0.1]) + rng.uniform(-0.15, 0.15, size = (200, 2)) * s
    random = rho0 * (1.0 + rng.uniform(-0.01, 0.01, size = len(jittered)))
    rate = diffusionRate(jittered, random, mode)
    volumes = rho0 * s * s / random
    assert abs(np.sum(volumes * rate)) < 1e-12 * np.sum(volumes * np.abs(rate))


# -- The density floor (tests/testSphSolver.py:765-783) -- #

@pytest.mark.parametrize('floor', [0.0, 900.0])
def testADensityFloorHoldsADropletWhoseNeighborsFlyApart(floor):
    # two particles of a droplet flying apart at 20 m/s for 20 steps lose most of their density
    # without a floor, and keep 0.9 rho0 with one
    positions = np.array([[0.0, 0.0], [s, 0.0]])
    p = physics(positions, 1.3 * s, 'wendlandC2', c0 = 50.0, densityFloor = floor, alpha = 0.0,
                domainMin = np.array([-1.0, -1.0]), domainMax = np.array([1.0, 1.0]))
    run = oracle.OracleRun(p, None, positions, np.array([[-10.0, 0.0], [10.0, 0.0]]), np.full(2, rho0),
                           np.full(2, rho0 * s * s), np.ones(2, dtype = bool))
    assert run.advance(np.inf, 20) == 20
    if floor:
        assert run.dens.min() == pytest.approx(900.0, rel = 1e-15)
    else:
        assert run.dens.min() < 900.0


# -- The periodic axis (tests/testSphSolver.py:793-816) -- #

def testAnExtrudedFlowInAPeriodicChannelStaysPlanar():
    # a random flow in (x, z), the same in every plane across a periodic y axis and drifting across it
    # at 3 m/s, keeps its spanwise velocity to rounding and wraps its particles back into the period
    # (the reference's flow has no drift, tests/testSphSolver.py:793-816; with one the planes cross
    # the period's end and must come back in at its start)
    drift = 3.0
    rng = np.random.default_rng(5)
    planes, period = 8, 8 * s
    section = np.array([(x, z) for x in np.arange(s / 2, 0.1, s) for z in np.arange(s / 2, 0.05, s)])
    sectionVelocity = rng.normal(0.0, 0.5, size = section.shape)
    across = (np.arange(planes) + 0.5) * s
    positions = np.column_stack([np.repeat(section[:, 0], planes), np.tile(across, len(section)),
                                 np.repeat(section[:, 1], planes)])
    velocities = np.column_stack([np.repeat(sectionVelocity[:, 0], planes), np.full(len(positions), drift),
                                  np.repeat(sectionVelocity[:, 1], planes)])
    count = len(positions)
    p = physics(positions, 1.3 * s, 'wendlandC2', period = period, domainMin = np.array([-1.0, 0.0, -1.0]),
                domainMax = np.array([1.0, period, 1.0]), alpha = 0.1)
    run = oracle.OracleRun(p, None, positions, velocities, np.full(count, rho0), np.full(count, rho0 * s**3),
                           np.ones(count, dtype = bool))
    assert run.advance(np.inf, 40) == 40
    assert np.abs(run.velocities[:, 1] - drift).max() < 1e-12
    assert np.abs(run.velocities[:, [0, 2]]).max() > 0.1
    assert drift * run.time.value > s
    assert np.all((run.positions[:, 1] >= 0.0) & (run.positions[:, 1] < period))


# -- Particle shifting (tests/testSphSolver.py:497-606) -- #

def shiftsOf(positions, velocities = None, densities = None, isFluid = None, size: float = 1e-5, h: float = 1.3 * s):
    count = len(positions)
    fluid = np.ones(count, dtype = bool) if isFluid is None else isFluid
    velocities = np.zeros((count, 2)) if velocities is None else velocities
    densities = np.full(count, rho0) if densities is None else densities
    p = physics(positions, h)
    return oracle.shiftingDisplacement(p, positions, velocities, densities, np.full(count, rho0 * s * s), fluid, size,
                                       False, s)


def testShiftingLeavesALatticeInsideTheWaterAlone():
    positions = lattice([0.0, 0.0], [0.2, 0.1])
    shift, _, _ = shiftsOf(positions)
    h = 1.3 * s
    top = positions[:, 1] > positions[:, 1].max() - 0.5 * s
    inside = np.all((positions > 3.0 * h) & (positions < np.array([0.2, 0.1]) - 3.0 * h), axis = 1)
    assert np.all(shift[top] == 0.0)
    assert np.abs(shift[inside]).max() < 1e-12


def testShiftingPartsACrowdedPair():
    positions = lattice([0.0, 0.0], [0.2, 0.2])
    centre = int(np.argmin(np.linalg.norm(positions - 0.1, axis = 1)))
    partner = int(np.argmin(np.linalg.norm(positions - (positions[centre] + [s, 0.0]), axis = 1)))
    positions[partner, 0] -= 0.4 * s
    shift, _, _ = shiftsOf(positions)
    assert shift[centre, 0] < 0.0 < shift[partner, 0]


def testShiftingNextToTheSurfaceRunsAlongIt():
    # Under the jittered flat top of a block of water, a particle with a surface particle among its
    # neighbors is shifted along the surface: its shift has no component along the kernel-weighted
    # mean of its surface neighbors' normals -grad C0_j / |grad C0_j| (surface: div r below 1.5, or
    # a lopsided support, h |grad C0| above 0.3 for the cubic spline)
    h = 1.3 * s
    rng = np.random.default_rng(2)
    positions = lattice([0.0, 0.0], [0.3, 0.1]) + rng.uniform(-0.1, 0.1, size = (300, 2)) * s
    shift, _, _ = shiftsOf(positions)
    tree = cKDTree(positions)
    pairs = np.array(sorted(tree.query_pairs(2.0 * h)))
    i, j = np.r_[pairs[:, 0], pairs[:, 1]], np.r_[pairs[:, 1], pairs[:, 0]]
    d = positions[i] - positions[j]
    r = np.linalg.norm(d, axis = 1)
    volume = s * s
    plain = np.zeros((len(positions), 2))
    np.add.at(plain, i, volume * kernelGradients('cubicSpline', d, r, h, 2))
    divergence = np.zeros(len(positions))
    np.add.at(divergence, i, -volume * kernelSlopes('cubicSpline', r, h, 2) * r)
    surface = (divergence < 1.5) | (h * np.linalg.norm(plain, axis = 1) > 0.3)
    own = -plain / np.linalg.norm(plain, axis = 1)[:, None]
    normal = np.zeros_like(plain)
    weight = kernelValues('cubicSpline', r, h, 2)
    np.add.at(normal, i, np.where(surface[j], weight, 0.0)[:, None] * own[j])
    vicinity = np.zeros(len(positions), dtype = bool)
    vicinity[i[surface[j]]] = True
    vicinity &= ~surface
    # most vicinity particles are held still by a steep concentration gradient (h |grad C| above 0.03)
    moved = vicinity & np.any(shift != 0.0, axis = 1)
    assert moved.sum() >= 1
    normal[moved] /= np.linalg.norm(normal[moved], axis = 1)[:, None]
    along = np.abs(np.sum(shift[moved] * normal[moved], axis = 1))
    assert along.max() < 1e-12 * np.abs(shift[moved]).max()
    # without the projection the shifts would cross the surface: grad C there has a normal part
    assert np.all(shift[surface] == 0.0)


def testAShiftCarriesTheDensityAndVelocityOfWhereItLands():
    # a shifted particle's density changes by shift . grad rho and its velocity by (shift . grad) v,
    # to the cubic spline's lattice error, for a linear density and a linear velocity field
    positions = lattice([0.0, 0.0], [0.2, 0.2])
    centre = int(np.argmin(np.linalg.norm(positions - 0.1, axis = 1)))
    positions[centre, 0] += 0.3 * s
    slope = np.array([3.0, -2.0])
    densities = rho0 + positions @ slope
    field = np.array([[0.4, -1.1], [0.7, 0.2]])
    shift, change, velocity = shiftsOf(positions, positions @ field.T, densities)
    # (the moved particle's lopsided support costs its difference gradients 4 percent)
    assert shift[centre, 0] < 0.0
    assert change[centre] == pytest.approx(np.dot(shift[centre], slope), rel = 0.05)
    assert velocity[centre] == pytest.approx(field @ shift[centre], rel = 0.05)
    # particles left where they are change nothing
    still = np.all(shift == 0.0, axis = 1)
    assert still.sum() > 0 and np.all(change[still] == 0.0) and np.all(velocity[still] == 0.0)


def testTheKernelNormalizationIsTheClosedForm():
    # the sigmas the restatement evaluates with: 10 / (7 pi h^2) and 7 / (4 pi h^2) in 2D,
    # 1 / (pi h^3) and 21 / (16 pi h^3) in 3D
    assert sigmaUnit('cubicSpline', 2) == pytest.approx(10.0 / (7.0 * math.pi), rel = 1e-15)
    assert sigmaUnit('wendlandC2', 2) == pytest.approx(7.0 / (4.0 * math.pi), rel = 1e-15)
    assert sigmaUnit('cubicSpline', 3) == pytest.approx(1.0 / math.pi, rel = 1e-15)
    assert sigmaUnit('wendlandC2', 3) == pytest.approx(21.0 / (16.0 * math.pi), rel = 1e-15)


# -- The equation of state and the integration of a step -- #

def tait(densities, soundSpeed: float = c0):
    '''Tait's p = rho0 c0^2 / 7 ((rho / rho0)^7 - 1).'''
    return rho0 * soundSpeed**2 / 7.0 * ((np.asarray(densities) / rho0)**7 - 1.0)


def testTheWallDensityCarriesTheExtrapolatedPressure():
    # a bed particle's density is the one Tait's equation turns into its extrapolated pressure,
    # rho0 (1 + p / B)^(1/7) (tests/testSphSolver.py:663-687 hold the pressure; this its density)
    water = lattice([0.0, 0.0], [0.3, 0.1])
    bed = lattice([0.0, -3.0 * s], [0.3, 0.0])
    positions = np.vstack([water, bed])
    fluid = np.arange(len(positions)) < len(water)
    densities = rho0 * (1.0 + 6.0 * 9.81 * (0.1 - positions[:, 1]) / c0**2)**(1.0 / 6.0)
    p = physics(positions, 1.3 * s, gravity = np.array([0.0, -9.81]))
    pressures, carried = oracle.wallPressure(p, positions, densities, tait(densities), np.full(len(positions), rho0 * s * s),
                                             fluid)
    reached = ~fluid & (pressures > 0.0)
    assert reached.sum() > 50
    assert tait(carried[reached]) == pytest.approx(pressures[reached], rel = 1e-10)
    assert np.all(carried[reached] > rho0)


def testTheSteppedPressureIsTaitsClosedForm():
    # the pressures a step of the restatement sums with: Tait's closed form at the step's densities,
    # clamped at zero on water below rho0 (the reference's default) and left negative with tension
    positions = lattice([0.0, 0.0], [0.1, 0.1])
    count = len(positions)
    densities = rho0 * (1.0 + np.random.default_rng(3).uniform(-0.01, 0.02, size = count))
    for tension in (False, True):
        p = physics(positions, 1.3 * s, negativePressure = tension, alpha = 0.0,
                    domainMin = np.array([-1.0, -1.0]), domainMax = np.array([1.0, 1.0]))
        run = oracle.OracleRun(p, None, positions, np.zeros((count, 2)), densities, np.full(count, rho0 * s * s),
                               np.ones(count, dtype = bool))
        run.advance(np.inf, 1)
        expected = tait(densities) if tension else np.maximum(tait(densities), 0.0)
        assert run.press == pytest.approx(expected, rel = 1e-12, abs = 1e-9)
        assert (run.press.min() < 0.0) == tension


def testAFreeParticleFallsAsTheSymplecticStep():
    # one particle alone under gravity at a fixed step: v_k+1 = v_k + g dt, then x_k+1 = x_k + v_k+1 dt
    # (the kick, then the drift with the kicked velocity), so after n steps
    # z = z0 + n w0 dt + g dt^2 n (n + 1) / 2 and x = x0 + n u0 dt
    dt, steps, g = 1e-4, 50, -9.81
    start, velocity = np.array([[0.2, 0.5]]), np.array([[0.3, 0.1]])
    p = physics(start, 1.3 * s, gravity = np.array([0.0, g]), maxTimeStep = dt,
                domainMin = np.array([-1.0, -1.0]), domainMax = np.array([1.0, 1.0]))
    run = oracle.OracleRun(p, None, start, velocity, np.full(1, rho0), np.full(1, rho0 * s * s), np.ones(1, dtype = bool))
    assert run.advance(np.inf, steps) == steps
    assert run.time.value == pytest.approx(steps * dt, rel = 1e-12)
    assert run.positions[0, 0] == pytest.approx(0.2 + steps * 0.3 * dt, rel = 1e-13)
    assert run.positions[0, 1] == pytest.approx(0.5 + steps * 0.1 * dt + g * dt**2 * steps * (steps + 1) / 2, rel = 1e-13)
    assert run.velocities[0] == pytest.approx([0.3, 0.1 + g * dt * steps], rel = 1e-13)
