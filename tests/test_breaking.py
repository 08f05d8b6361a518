'''

The breaking read-out (pysurf_amd/breaking.py) on constructed particle arrangements whose answers
are fixed by construction, the arrangements of tests/testWaveBreaking.py:30-198: a face sloping back
leans negatively, a vertical face not at all, an overhanging lip by its overhang; still water ahead,
a front running past the window and rows lifted under the crest are not lean; the tracker records
breaking only once the lean persists, at the first update of the streak; the profile, trough and
waterline are the arrangement's.

'''

import numpy as np
import pytest

from pysurf_amd.bathymetry import Bathymetry
from pysurf_amd.breaking import BreakingTracker, freeSurfaceProfile, frontFaceLean, troughAhead, waterlineX
from pysurf_amd.particles import fillRegion, latticePoints

s = 0.05
swl = 1.0


def column(x0, x1, z0, z1):
    return latticePoints([x0, z0], [x1, z1], s)


def crestWithFace(faceX):
    '''Still water to z = 1 over -1 < x < 3, and a crest from x = 0 whose face at height z stands at faceX(z).'''
    water = column(-1.0, 3.0, 0.0, swl)
    crest = fillRegion(lambda p: (p[:, 0] > 0.0) & (p[:, 0] < faceX(p[:, 1])), [-1.0, swl], [3.0, swl + 0.4], s)
    return np.vstack([water, crest])


def testAFaceSlopingBackHasNegativeLean():
    assert frontFaceLean(crestWithFace(lambda z: 1.0 - 2.0 * (z - swl)), 0.0, swl, s, reach = 2.0) < -s


def testAVerticalFaceHasNoLean():
    assert frontFaceLean(crestWithFace(lambda z: np.full_like(z, 1.0)), 0.0, swl, s, reach = 2.0) == \
        pytest.approx(0.0, abs = 1e-12)


def testALipLeansByItsOverhang():
    positions = crestWithFace(lambda z: np.where(z < swl + 0.2, 1.0, 1.3))
    assert frontFaceLean(positions, 0.0, swl, s, reach = 2.0) == pytest.approx(0.3, abs = 1e-9)


def testStillWaterAheadDoesNotMaskTheFace():
    positions = crestWithFace(lambda z: np.full_like(z, 1.0))
    assert np.max(positions[:, 0]) > 2.9
    assert frontFaceLean(positions, 0.0, swl, s, reach = 2.0) <= 0.0


def testAFaceRunningPastTheWindowIsNotLean():
    rng = np.random.default_rng(5)
    positions = crestWithFace(lambda z: 4.0 - 8.0 * (z - swl))
    positions += rng.uniform(-0.3 * s, 0.3 * s, size = positions.shape)
    assert frontFaceLean(positions, 0.0, swl, s, reach = 1.0) < 0.5 * s


def testRowsLiftedUnderTheCrestAreNotLean():
    for seed in range(20):
        positions = latticePoints([-3.0, 0.0], [4.0, swl], s)
        positions[:, 1] *= 1.0 + 0.3 / np.cosh(4.0 * positions[:, 0])**2 / swl
        positions += np.random.default_rng(seed).uniform(-0.3 * s, 0.3 * s, size = positions.shape)
        assert frontFaceLean(positions, 0.0, swl, s, reach = 1.0) < s


def testTooFewBandsGiveMinusInfinity():
    assert frontFaceLean(column(-1.0, 3.0, 0.0, swl), 0.0, swl, s, reach = 2.0) == -np.inf


def testTheProfileIsTheTopOfTheWater():
    positions = np.vstack([column(0.0, 1.0, 0.0, 0.5), column(1.0, 2.0, 0.0, 0.8)])
    centers, elevation = freeSurfaceProfile(positions, 0.0, 3.0, s)
    assert centers[0] == pytest.approx(0.5 * s)
    assert elevation[centers < 1.0] == pytest.approx(np.full(np.sum(centers < 1.0), 0.5))
    assert elevation[(centers > 1.0) & (centers < 2.0)] == pytest.approx(np.full(20, 0.8))
    assert np.all(np.isnan(elevation[centers > 2.0]))


def testTheWaterlineIsTheTipOfTheSheet():
    # A sheet of still water at 0.5 on a 1:10 beach meets the bed at x = 5; a droplet ahead does not count
    beach = Bathymetry([0.0, 10.0], [0.0, 1.0], stillWaterLevel = 0.5)
    sheet = fillRegion(lambda p: (p[:, 1] > beach.elevation(p[:, 0])) & (p[:, 1] < 0.5), [0.0, 0.0], [10.0, 0.6], s)
    tip = waterlineX(np.vstack([sheet, np.array([[8.0, 0.82]])]), beach, wetHeight = 2.0 * s)
    assert tip == pytest.approx(5.0, abs = 2.0 * s / 0.1)
    assert tip < 6.0


def testTheTroughIsTheLowestWaterAheadOfTheCrest():
    centers = np.arange(12) + 0.5
    elevation = np.array([0.0, 0.3, 0.1, -0.1, -0.1, 0.0, np.nan, -0.5, -0.5, 0.0, 0.0, 0.0])
    assert troughAhead(centers, elevation, 1.5) == pytest.approx(-0.1)
    assert np.isnan(troughAhead(centers, elevation, 5.5))
    assert np.isnan(troughAhead(centers, elevation, 11.5))
    still = np.zeros(12)
    still[7] = -0.12
    still[11] = -0.05
    assert troughAhead(centers, still, 1.5) == pytest.approx(0.0)
    assert troughAhead(centers, elevation, 0.5, reach = 3.5) == pytest.approx(0.1)


def testTheTrackerRecordsTheCrestAndBreaking():
    beach = Bathymetry([-1.0, 3.0, 20.0], [0.0, 0.0, 1.7], stillWaterLevel = swl)
    tracker = BreakingTracker(beach, s, faceReach = 2.0, searchFromX = -0.5)
    vertical = crestWithFace(lambda z: np.full_like(z, 1.0))
    lip = crestWithFace(lambda z: np.where(z < swl + 0.2, 1.0, 1.3))
    tracker.update(0.0, vertical)
    assert tracker.breakingPoint is None
    # A lip seen once, then gone, is jitter: the streak starts again
    tracker.update(0.1, lip)
    tracker.update(0.2, vertical)
    for time in (0.3, 0.4):
        tracker.update(time, lip)
    assert tracker.breakingPoint is None
    tracker.update(0.5, lip)
    point = tracker.breakingPoint
    assert point is not None and point.time == 0.3
    assert point.lean == pytest.approx(0.3, abs = 1e-9)
    assert point.crestHeight == pytest.approx(0.4, abs = s)
    assert point.breakingIndex == pytest.approx(point.crestHeight / point.depth)
    assert point.troughHeight == pytest.approx(0.0, abs = 1e-9)
    assert point.waveHeight == pytest.approx(point.crestHeight, abs = 1e-9)
    assert len(tracker.times) == 6


def testTheTrackerReadsTheDrawdownAheadOfABreakingCrest():
    beach = Bathymetry([-1.0, 3.0, 20.0], [0.0, 0.0, 1.7], stillWaterLevel = swl)
    tracker = BreakingTracker(beach, s, faceReach = 2.0, searchFromX = -0.5)
    lowered = np.vstack([column(-1.0, 1.5, 0.0, swl), column(1.5, 3.0, 0.0, swl - 0.1),
                         fillRegion(lambda p: (p[:, 0] > 0.0) & (p[:, 0] < np.where(p[:, 1] < swl + 0.2, 1.0, 1.3)),
                                    [-1.0, swl], [3.0, swl + 0.4], s)])
    for time in (0.0, 0.1, 0.2):
        tracker.update(time, lowered)
    point = tracker.breakingPoint
    assert point is not None
    assert point.troughHeight == pytest.approx(-0.1, abs = 1e-9)
    assert point.waveHeight == pytest.approx(point.crestHeight + 0.1, abs = 1e-9)


def testAnUnknownFaceMeasureIsRefused():
    with pytest.raises(ValueError):
        BreakingTracker(Bathymetry([0.0, 1.0], [0.0, 0.0], 1.0), s, faceMeasure = 'bands')
