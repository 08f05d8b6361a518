'''

The host half of the device read-out, without a GPU: the waterline decided from the most landward
candidates is the waterline of the full set whenever it is decided, and the occupancy words unpack
to the face window's cells in the layout sph_face_cells writes.

'''

import numpy as np

from pysurf_amd.bathymetry import Bathymetry
from pysurf_amd.breaking import faceCells, waterlineFromLandward, waterlineX
from pysurf_amd.readout import LANDWARD, occupancyRows, trackedPositions


def landwardOf(points: np.ndarray, bed, wetHeight: float) -> tuple[np.ndarray, bool]:
    '''What sph_face_cells returns for one run: the LANDWARD most landward points of the widened set, NaN-padded.'''
    widened = points[points[:, 1] - bed.elevation(points[:, 0]) < wetHeight + 1e-9]
    ordered = widened[np.argsort(-widened[:, 0], kind = 'stable')][:LANDWARD]
    landward = np.full((LANDWARD, 2), np.nan)
    landward[:len(ordered)] = ordered
    return landward, len(ordered) < LANDWARD


def testTheLandwardWaterlineIsTheFullWaterline():
    bed = Bathymetry.planeBeach(depth = 1.0, toeX = 2.0, slope = 1.0 / 15.0, endX = 30.0)
    wet = 0.25
    rng = np.random.default_rng(3)
    decided = 0
    for trial in range(400):
        count = int(rng.integers(0, 40))
        x = rng.uniform(0.0, 30.0, count)
        # Heights spread over the band's edge, some exactly on it and some a rounding step either side
        height = rng.uniform(0.0, 2.0 * wet, count)
        edge = rng.random(count) < 0.2
        height[edge] = rng.choice([wet, np.nextafter(wet, 0.0), wet + 5e-10, wet - 5e-10], int(edge.sum()))
        points = np.stack([x, bed.elevation(x) + height], axis = 1)
        landward, complete = landwardOf(points, bed, wet)
        tip = waterlineFromLandward(landward, bed, wet, complete)
        exact = waterlineX(points, bed, wet)
        if tip is None:
            assert not complete
            continue
        decided += 1
        assert tip == exact or (np.isnan(tip) and np.isnan(exact)), (trial, tip, exact)
    assert decided > 350


def testAnUndecidedWaterlineAsksForThePositions():
    # Eight candidates just above the wet band, closer to shore than the three that are in it
    bed = Bathymetry.planeBeach(depth = 1.0, toeX = 2.0, slope = 1.0 / 15.0, endX = 30.0)
    wet = 0.25
    x = np.concatenate([np.linspace(20.0, 21.0, 8), [10.0, 11.0, 12.0]])
    height = np.concatenate([np.full(8, wet + 5e-10), np.full(3, 0.1)])
    points = np.stack([x, bed.elevation(x) + height], axis = 1)
    landward, complete = landwardOf(points, bed, wet)
    assert not complete
    assert waterlineFromLandward(landward, bed, wet, complete) is None
    assert waterlineX(points, bed, wet) == 10.0


def testOccupancyWordsUnpackToTheFaceCells():
    rng = np.random.default_rng(5)
    positions = np.stack([rng.uniform(3.0, 6.0, 3000), rng.uniform(0.9, 1.6, 3000)], axis = 1)
    occupied, _ = faceCells(positions, 4.0, 1.0, 0.05, 1.7)
    bands, columns = occupied.shape
    assert columns > 32 and columns % 32 != 0
    perRow = -(-columns // 32)
    offset = 7
    words = np.zeros(offset + bands * perRow, dtype = np.uint32)
    for band, column in zip(*np.nonzero(occupied)):
        words[offset + band * perRow + column // 32] |= np.uint32(1) << np.uint32(column % 32)
    assert np.array_equal(occupancyRows(words, offset, bands, columns), occupied)


def testTheTrackedStripIsTheCentreLine():
    positions = np.array([[1.0, 0.0, 0.5], [2.0, 0.2, 0.6], [3.0, -0.1, 0.7], [4.0, 0.05, 0.8]])
    fluid = np.array([True, True, True, False])
    assert np.array_equal(trackedPositions(positions, fluid, 3, 0.0, 0.125), [[1.0, 0.5], [3.0, 0.7]])
    assert np.array_equal(trackedPositions(positions[:, :2], fluid, 2), positions[:3, :2])
