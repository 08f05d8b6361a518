'''

The smoothed-surface read-out ('contour', pysurf_amd/breaking.py) on the arrangements of
tests/testWaveBreaking.py:290-530, whose answers are fixed by construction: the surface stands at
the top of the water's cells and its closed contours sort into pockets of air and drops of water;
the face reads level, vertical or overhanging as built, over two or six spacings of arc, in deep or
shallow water, whatever voids lie under the crest; the touchdown is the tube closing under a lip,
not a void opening in the crest nor a pocket under a face that has sloped back; the breaking point
is the start of the last vertical run before touchdown, or of the longest without one; split
particles count their own area.

The fluid fraction itself is pinned by what the kernel fixes: it is one inside a lattice and one
half on a lattice's edge, the nodes' sum times the node area is the particles' area, and a single
particle's value at a node is its area times the 2D cubic spline there.

'''

import math

import numpy as np
import pytest

from pysurf_amd.bathymetry import Bathymetry
from pysurf_amd.breaking import (BreakingTracker, ParticleSurface, airShare, contourFace, cubicSpline2D, fluidFraction,
                                 frontFaceLean, particleSurface)
from pysurf_amd.particles import ParticleSystem, fillRegion, latticePoints
from pysurf_amd.refinement import splitParticles

s = 0.05
swl = 1.0
flatBed = Bathymetry([-3.0, 4.0], [0.0, 0.0], stillWaterLevel = swl)


def column(x0, x1, z0, z1):
    return latticePoints([x0, z0], [x1, z1], s)


def crestWithFace(faceX):
    '''Still water to z = 1 over -1 < x < 3, and a crest from x = 0 whose face at height z stands at faceX(z).'''
    water = column(-1.0, 3.0, 0.0, swl)
    crest = fillRegion(lambda p: (p[:, 0] > 0.0) & (p[:, 0] < faceX(p[:, 1])), [-1.0, swl], [3.0, swl + 0.4], s)
    return np.vstack([water, crest])


sloping = lambda z: 1.0 - 0.5 * (z - swl)
vertical = lambda z: np.full_like(z, 1.0)
lipped = lambda z: np.where(z < swl + 0.2, 1.0, 1.3)


# -- The fluid fraction -- #

def testTheSplineIsTheTextbookCubic():
    h = 0.065
    sigma = 10.0 / (7.0 * math.pi * h * h)
    assert cubicSpline2D(np.array([0.0]), h)[0] == pytest.approx(sigma)
    assert cubicSpline2D(np.array([h]), h)[0] == pytest.approx(0.25 * sigma)
    assert cubicSpline2D(np.array([1.5 * h]), h)[0] == pytest.approx(sigma * 0.25 * 0.125)
    assert cubicSpline2D(np.array([2.0 * h, 3.0 * h]), h).tolist() == [0.0, 0.0]
    # Its integral over the plane is one
    r = np.linspace(0.0, 2.0 * h, 20001)
    assert np.trapezoid(2.0 * math.pi * r * cubicSpline2D(r, h), r) == pytest.approx(1.0, rel = 1e-6)


def testTheFractionIsOneInsideAndOneHalfOnTheEdge():
    water = column(0.0, 2.0, 0.0, 1.0)
    gridX, gridY, phi = fluidFraction(water, s, np.array([0.5, 0.5]), np.array([1.5, 1.2]))
    inside = phi[np.ix_(gridY < 0.85, (gridX > 0.6) & (gridX < 1.4))]
    assert np.abs(inside - 1.0).max() < 0.01
    edge = phi[np.argmin(np.abs(gridY - 1.0)), (gridX > 0.6) & (gridX < 1.4)]
    assert np.abs(edge - 0.5).max() < 0.01
    assert phi[gridY > 1.0 + 2.0 * 1.3 * s].max() == 0.0


def testTheFractionOfOneParticleIsItsKernel():
    point = np.array([[0.313, 0.271]])
    gridX, gridY, phi = fluidFraction(point, s, np.array([0.0, 0.0]), np.array([0.6, 0.6]), volumes = np.array([0.7]))
    nodeX, nodeY = np.meshgrid(gridX, gridY)
    expected = 0.7 * cubicSpline2D(np.hypot(nodeX - point[0, 0], nodeY - point[0, 1]), 1.3 * s)
    assert np.allclose(phi, expected, rtol = 0.0, atol = 1e-12)
    # The nodes' sum over the node area is the particle's area
    assert phi.sum() * (0.5 * s)**2 == pytest.approx(0.7, rel = 1e-3)


# -- The surface, pockets and drops (tests/testWaveBreaking.py:290-309) -- #

def testTheParticleSurfaceSortsTheSurfacePocketsAndDrops():
    water = column(0.0, 1.0, 0.0, 0.5)
    hole = (np.abs(water[:, 0] - 0.5) < 0.2 - 1e-9) & (np.abs(water[:, 1] - 0.25) < 0.1 - 1e-9)
    drop = column(0.4, 0.55, 0.7, 0.85)
    result = particleSurface(np.vstack([water[~hole], drop]), s, np.array([-0.2, 0.1]), np.array([1.2, 1.0]))
    assert result.surface[0, 0] < result.surface[-1, 0]
    top = result.surface[np.abs(result.surface[:, 0] - 0.5) < 0.3]
    assert np.allclose(top[:, 1], 0.5, atol = 0.05 * s)
    assert len(result.pockets) == 1 and len(result.drops) == 1
    pocket, droplet = result.pockets[0], result.drops[0]
    assert np.allclose([pocket[:, 0].min(), pocket[:, 0].max(), pocket[:, 1].min(), pocket[:, 1].max()],
                       [0.3, 0.7, 0.15, 0.35], atol = 0.05 * s)
    assert ParticleSurface.area(pocket) == pytest.approx(0.4 * 0.2, rel = 0.1)
    assert ParticleSurface.area(droplet) == pytest.approx(0.15 * 0.15, rel = 0.25)


def testTheShoelaceAreaOfASquare():
    square = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0], [0.0, 2.0], [0.0, 0.0]])
    assert ParticleSurface.area(square) == 4.0
    assert ParticleSurface.area(square[::-1]) == 4.0


# -- The face from the surface (tests/testWaveBreaking.py:311-403) -- #

@pytest.mark.parametrize('face, slope', [(sloping, 0.5 / np.sqrt(1.25)), (vertical, 0.0), (lipped, -1.0)])
def testTheContourFaceReadsTheSlopeOfTheFace(face, slope):
    result = contourFace(crestWithFace(face), s, 0.5, flatBed, reach = 1.0, windowDepth = 0.1)
    assert result.crestHeight == pytest.approx(0.4, abs = 0.05 * s)
    assert result.slope == pytest.approx(slope, abs = 0.02)
    assert 0.75 < result.faceX < 1.35


def testALongerArcReadsTheFaceAndNotAStepInIt():
    plane = crestWithFace(sloping)
    for arc in (2.0, 6.0):
        result = contourFace(plane, s, 0.5, flatBed, reach = 1.0, windowDepth = 0.1, arcLength = arc)
        assert result.slope == pytest.approx(0.5 / np.sqrt(1.25), abs = 0.02)
    low, high = swl + 0.15, swl + 0.15 + 3.0 * s
    held = 1.0 - 0.6 * (low - swl)

    def stepped(z):
        return np.where(z < low, 1.0 - 0.6 * (z - swl), np.where(z < high, held, held - 0.6 * (z - high)))

    positions = crestWithFace(stepped)
    short = contourFace(positions, s, 0.5, flatBed, reach = 1.0, windowDepth = 0.1, arcLength = 2.0)
    long = contourFace(positions, s, 0.5, flatBed, reach = 1.0, windowDepth = 0.1, arcLength = 6.0)
    assert abs(short.slope) < 0.05
    assert long.slope > 0.2
    assert short.faceX == pytest.approx(held, abs = 2.0 * s) and long.faceX == pytest.approx(held, abs = 2.0 * s)
    for arc, check in ((2.0, lambda slope: abs(slope) < 0.05), (6.0, lambda slope: slope > 0.2)):
        tracker = BreakingTracker(flatBed, s, faceReach = 1.0, faceMeasure = 'contour', faceArc = arc)
        tracker.update(1.0, positions)
        assert check(tracker.faceSlope[-1])


def testTheContourFaceReadsAFaceInShallowWater():
    shallowBed = Bathymetry([-3.0, 4.0], [0.7, 0.7], stillWaterLevel = swl)
    water = column(-1.0, 3.0, 0.7, swl)
    crest = fillRegion(lambda p: (p[:, 0] > 0.0) & (p[:, 0] < 1.0), [-1.0, swl], [3.0, swl + 0.4], s)
    result = contourFace(np.vstack([water, crest]), s, 0.5, shallowBed, reach = 1.0, windowDepth = 0.2)
    assert result.window is not None
    assert result.crestHeight == pytest.approx(0.4, abs = 0.05 * s)
    assert result.slope == pytest.approx(0.0, abs = 0.02)
    # Water 0.24 deep leaves a window 0.045 deep, less than a spacing: there is none
    assert contourFace(np.vstack([water, crest]), s, 0.5, Bathymetry([-3.0, 4.0], [0.76, 0.76], stillWaterLevel = swl),
                       reach = 1.0, windowDepth = 0.2).window is None


def testAVoidUnderTheCrestIsNotAFace():
    positions = crestWithFace(sloping)
    void = (positions[:, 1] > swl + 0.1) & (positions[:, 1] < swl + 0.2) & (positions[:, 0] > 0.1) & (positions[:, 0] < 0.6)
    positions = positions[~void]
    assert frontFaceLean(positions, 0.0, swl, s, reach = 1.0) > 10.0 * s
    assert contourFace(positions, s, 0.5, flatBed, reach = 1.0, windowDepth = 0.1).slope == pytest.approx(0.447, abs = 0.02)
    assert len(particleSurface(positions, s, np.array([-0.5, 0.9]), np.array([2.0, 2.0])).pockets) == 1


# -- Touchdown and the breaking point (tests/testWaveBreaking.py:405-514) -- #

def testTheTrackerRecordsTouchdownWhenTheTubeCloses():
    lip = crestWithFace(lipped)
    tip = column(1.25, 1.3, swl, swl + 0.2)
    tracker = BreakingTracker(flatBed, s, faceReach = 1.0, faceMeasure = 'contour')
    tracker.update(1.0, lip)
    assert tracker.touchdownSurface is None
    tracker.update(2.0, np.vstack([lip, tip]))
    assert tracker.touchdownTime == pytest.approx(1.5)
    assert np.any(np.diff(tracker.touchdownSurface[:, 0]) < -0.5 * s)
    assert ParticleSurface.area(tracker.touchdownPocket) == pytest.approx(0.25 * 0.2, rel = 0.15)


def testAVoidInTheWaterIsNotATouchdown():
    lip = crestWithFace(lipped)
    inVoid = (lip[:, 0] > 0.1) & (lip[:, 0] < 0.5) & (lip[:, 1] > swl + 0.25) & (lip[:, 1] < swl + 0.35)
    voided = lip[~inVoid]
    tip = column(1.25, 1.3, swl, swl + 0.2)
    tracker = BreakingTracker(flatBed, s, faceReach = 1.0, faceMeasure = 'contour')
    tracker.update(1.0, lip)
    tracker.update(2.0, voided)
    assert tracker.touchdownSurface is None
    window = particleSurface(voided, s, np.array([-0.5, 0.9]), np.array([2.0, 2.0]))
    assert len(window.pockets) == 1 and airShare(window.pockets[0], tracker.lastSurface, s) == 0.0
    tracker.update(3.0, np.vstack([voided, tip]))
    assert tracker.touchdownTime == pytest.approx(2.5)
    assert 1.0 < tracker.touchdownPocket[:, 0].mean() < 1.25


def testTheBreakingPointIsTheLastVerticalRunBeforeTouchdown():
    slope, lip = crestWithFace(sloping), crestWithFace(lipped)
    tip = column(1.25, 1.3, swl, swl + 0.2)
    tracker = BreakingTracker(flatBed, s, faceReach = 1.0, faceMeasure = 'contour')
    tracker.update(0.0, slope)
    assert tracker.breakingPoint is None
    for time in (1.0, 2.0, 3.0):
        tracker.update(time, lip)
    tracker.update(4.0, slope)
    tracker.update(5.0, lip)
    assert tracker.breakingPoint.time == 1.0 and not tracker.breakingFinal
    tracker.update(6.0, np.vstack([lip, tip]))
    tracker.update(7.0, slope)
    tracker.update(8.0, lip)
    assert tracker.breakingFinal and tracker.touchdownTime == pytest.approx(5.5)
    assert tracker.breakingPoint.time == 5.0
    assert 0.75 < tracker.breakingPoint.faceX < 1.35
    assert tracker.breakingPoint.faceDepth == pytest.approx(swl)


def testWithoutTouchdownTheBreakingPointIsTheLongestVerticalRun():
    slope, lip = crestWithFace(sloping), crestWithFace(lipped)
    tracker = BreakingTracker(flatBed, s, faceReach = 1.0, faceMeasure = 'contour')
    for time, face in enumerate([lip, slope, lip, lip, lip, slope, lip, slope]):
        tracker.update(float(time), face)
    assert tracker.breakingPoint.time == 2.0 and not tracker.breakingFinal


def testATubeUnderAFaceThatSlopedBackIsNotATouchdown():
    slope, lip = crestWithFace(sloping), crestWithFace(lipped)
    tip = column(1.25, 1.3, swl, swl + 0.2)
    tracker = BreakingTracker(flatBed, s, faceReach = 1.0, faceMeasure = 'contour')
    tracker.update(0.0, lip)
    tracker.update(1.0, slope)
    tracker.update(2.0, np.vstack([lip, tip]))
    assert tracker.touchdownSurface is None and not tracker.breakingFinal


def testACrestOffshoreOfTheSearchIsNotRead():
    tracker = BreakingTracker(flatBed, s, faceReach = 1.0, faceMeasure = 'contour', searchFromX = 2.0)
    tracker.update(0.0, crestWithFace(lipped))
    assert np.isnan(tracker.faceSlope[-1]) and tracker.breakingPoint is None
    assert 'grid' not in tracker.measures(crestWithFace(lipped))


def testSplitParticlesCountTheirOwnAreaInTheSurface():
    spacing = 0.02
    particles = ParticleSystem.assemble(latticePoints(np.zeros(2), np.array([1.0, 0.5]), spacing), np.empty((0, 2)),
                                        spacing, 1000.0)
    splitParticles(particles, particles.positions[:, 0] > 0.5, 1000.0, 1.3 * spacing)
    lower, upper = np.array([0.1, 0.3]), np.array([0.9, 0.7])
    weighted = particleSurface(particles.positions, spacing, lower, upper, volumes = particles.masses / 1000.0).surface
    counted = particleSurface(particles.positions, spacing, lower, upper).surface
    assert np.abs(weighted[:, 1] - 0.5).max() < 0.1 * spacing
    assert counted[counted[:, 0] > 0.6, 1].min() > 0.5 + 0.5 * spacing
