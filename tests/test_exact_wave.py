'''

The exact solitary wave and the δ-SPH start laid from it (pysurf_amd/exact_wave.py), against the
published values and the reference's checks (tests/testWaveTheory.py:271-303,
tests/testWaveMakers.py:226-281, tests/testSphParticles.py:45-90), and against what potential flow
fixes: no pressure on the free surface, no flow through the bed, no divergence, still water far
away. The Grilli 1:35 start holds the reference's particle count (128,947 fluid particles at 64 per
depth, validation/gpuSph/results/grilliBreaking_64.json).

'''

import math

import numpy as np
import pytest

from pysurf_amd.bathymetry import Bathymetry
from pysurf_amd.exact_wave import (babenkoSolitaryWave, bedFittedPoints, exactSolitaryBoundaryPressure,
                                   exactSolitaryFlow, exactSolitaryInterior, exactSolitarySurface, exactSolitaryWave,
                                   exactSolitaryWaveFluid, exactSurfaceElevation)
from pysurf_amd.particles import latticePoints

g = 9.80665


def testTheExactSolitaryWaveMatchesClamondAndDutykh():
    # Clamond & Dutykh (2013), table 1: F = 1.274236297977903 gives a/d = 0.68196566738690
    assert babenkoSolitaryWave(1.274236297977903).amplitude == pytest.approx(0.68196566738690, abs = 5e-12)


def testTheExactSolitaryWaveSpeedMatchesGuyenneAndGrilli():
    # Guyenne & Grilli (2006): Tanaka's wave of H/d = 0.6 travels at c / sqrt(g d) = 1.25
    assert exactSolitaryWave(0.6).froudeNumber == pytest.approx(1.25, abs = 5e-3)


def testASmallWaveIsTheKdVWave():
    # At H/d = 0.05 the speed is the first-order sqrt(1 + H/d) to O((H/d)^2), the shape sech^2 to O(H/d)
    wave = exactSolitaryWave(0.05)
    assert wave.froudeNumber == pytest.approx(math.sqrt(1.05), abs = 2e-3)
    kdv = 0.05 / np.cosh(math.sqrt(0.75 * 0.05) * wave.x)**2
    assert np.abs(wave.eta - kdv).max() < 0.05 * 0.05


def testTheExactSolitaryWaveIsSymmetricAndSingleValued():
    wave = exactSolitaryWave(0.6)
    crest = len(wave.x) // 2
    assert np.all(np.diff(wave.x) > 0.0)
    assert wave.x[crest] == 0.0 and wave.potential[crest] == 0.0
    assert wave.eta[crest + 1:crest + 200] == pytest.approx(wave.eta[crest - 1:crest - 200:-1], abs = 1e-12)
    assert wave.potential[crest + 1:crest + 200] == pytest.approx(-wave.potential[crest - 1:crest - 200:-1], abs = 1e-12)


def testTheExactSolitarySurfaceIsSampledAtTheCrestAndStartsAtZeroPotential():
    x = np.linspace(0.0, 20.0, 2001)
    eta, potential = exactSolitarySurface(x, 0.6, 1.0, crestX = 8.0)
    assert eta.max() == pytest.approx(0.6, abs = 1e-6) and x[np.argmax(eta)] == pytest.approx(8.0)
    assert potential[0] == 0.0 and np.all(np.diff(potential) > 0.0)


def testTheFlowMeetsTheSurfaceAndBedConditions():
    x, y, u, v = exactSolitaryInterior(0.3)
    # No flow through the bed, which lies at y = -1
    assert np.abs(y[0] + 1.0).max() < 1e-12 and np.abs(v[0]).max() < 1e-12
    # The surface level is the wave
    wave = exactSolitaryWave(0.3)
    assert np.interp(0.0, x[-1], y[-1]) == pytest.approx(0.3, abs = 1e-9)
    # No pressure on the free surface: Bernoulli in the wave's frame holds there
    surface = np.column_stack([x[-1][::64], y[-1][::64]])
    _, pressure, _ = exactSolitaryFlow(surface - [0.0, 1e-12], 0.3, 1.0, 0.0, 0.0)
    assert np.abs(pressure).max() < 1e-6 * g
    # Still water far from the crest
    velocity, pressure, elevation = exactSolitaryFlow(np.array([[60.0, -0.5], [-60.0, -0.25]]), 0.3, 1.0)
    assert np.abs(velocity).max() < 1e-9 and np.abs(elevation).max() < 1e-9
    assert pressure == pytest.approx([0.5 * g, 0.25 * g], rel = 1e-9)


def testTheFlowIsDivergenceFree():
    step = 1e-4
    points = np.array([[0.3, -0.5], [-1.0, -0.2], [1.5, 0.0], [0.0, -0.9]])
    def velocity(offset):
        return exactSolitaryFlow(points + offset, 0.3, 1.0)[0]
    dudx = (velocity([step, 0.0])[:, 0] - velocity([-step, 0.0])[:, 0]) / (2.0 * step)
    dvdy = (velocity([0.0, step])[:, 1] - velocity([0.0, -step])[:, 1]) / (2.0 * step)
    speed = np.abs(velocity([0.0, 0.0])).max()
    assert np.abs(dudx + dvdy).max() < 0.02 * speed


def testBedFittedColumnsStandHalfASpacingAboveASlopingFloor():
    spacing = 0.05
    points = bedFittedPoints([0.0, 0.0], [1.0, 1.0], spacing, lambda x: 0.1 + 0.3 * x)
    columns = np.unique(points[:, 0])
    assert columns == pytest.approx(np.arange(0.025, 1.0, spacing))
    for x in columns:
        z = np.sort(points[points[:, 0] == x, 1])
        assert z[0] == pytest.approx(0.1 + 0.3 * x + 0.5 * spacing)
        assert np.diff(z) == pytest.approx(spacing)
        assert 1.0 - spacing <= z[-1] < 1.0


@pytest.mark.parametrize('lower, upper', [([0.0, 0.0], [1.0, 0.5]), ([0.0, 0.0, 0.0], [0.2, 0.3, 0.4])])
def testBedFittedColumnsOnAFlatFloorAreTheLattice(lower, upper):
    ordered = lambda points: points[np.lexsort(points.T[::-1])]
    fitted = bedFittedPoints(lower, upper, 0.1, lambda *horizontal: np.zeros_like(horizontal[0]))
    lattice = latticePoints(lower, upper, 0.1)
    assert fitted.shape == lattice.shape
    assert ordered(fitted) == pytest.approx(ordered(lattice), abs = 1e-12)


def testABedFittedExactWaveStandsHalfASpacingAboveTheSlope():
    bed = Bathymetry([0.0, 4.0, 10.0], [0.0, 0.0, 0.4], stillWaterLevel = 1.0)
    spacing = 1.0 / 16.0
    fitted = exactSolitaryWaveFluid(bed, 0.3, 2.5, spacing, bedFitted = True)
    lattice = exactSolitaryWaveFluid(bed, 0.3, 2.5, spacing)

    def gaps(points):
        columns, inverse = np.unique(points[:, 0], return_inverse = True)
        bottom = np.full(columns.size, np.inf)
        np.minimum.at(bottom, inverse, points[:, 1])
        return bottom - bed.elevation(columns)

    assert gaps(fitted[0]) == pytest.approx(0.5 * spacing)
    latticeGap = gaps(lattice[0])
    assert latticeGap.max() > spacing and latticeGap.min() >= 0.5 * spacing - 1e-9

    def flatPart(start):
        positions, velocities, pressures = start
        keep = positions[:, 0] < 4.0
        order = np.lexsort(positions[keep].T[::-1])
        return [positions[keep][order], velocities[keep][order], pressures[keep][order]]

    for fittedPart, latticePart in zip(flatPart(fitted), flatPart(lattice)):
        assert fittedPart == pytest.approx(latticePart, abs = 1e-12)
    assert len(fitted[0]) > len(lattice[0])


def testAJitteredExactWaveKeepsItsWaterAndCarriesTheFlowWhereItLies():
    bed = Bathymetry([0.0, 4.0, 10.0], [0.0, 0.0, 0.4], stillWaterLevel = 1.0)
    spacing = 1.0 / 16.0
    fitted, _, _ = exactSolitaryWaveFluid(bed, 0.3, 2.5, spacing, bedFitted = True)
    positions, velocities, pressures = exactSolitaryWaveFluid(bed, 0.3, 2.5, spacing, bedFitted = True, jitter = 0.15,
                                                              seed = 3)
    flow, pressure, surface = exactSolitaryFlow(positions, 0.3, 1.0, 2.5, 1.0)
    assert len(positions) == len(fitted)
    assert np.count_nonzero(np.all(positions == fitted, axis = 1)) < 0.01 * len(fitted)
    assert np.all(positions[:, 1] >= bed.elevation(positions[:, 0]) + 0.5 * spacing - 1e-12)
    assert np.all(positions[:, 1] < 1.0 + surface)
    assert exactSurfaceElevation(positions[:, 0], 0.3, 1.0, 2.5) == pytest.approx(surface, abs = 1e-12)
    assert velocities == pytest.approx(flow, abs = 1e-12)
    assert pressures == pytest.approx(np.maximum(pressure, 0.0), abs = 1e-12)


def testTheWallsStartWithTheWaterPressureJustInside():
    bed = Bathymetry([0.0, 4.0, 10.0], [0.0, 0.0, 0.4], stillWaterLevel = 1.0)
    spacing = 1.0 / 16.0
    walls = np.array([[50.0, -0.5 * spacing], [1.0, -1.5 * spacing], [-0.5 * spacing, 0.5], [2.5, 2.0]])
    pressure = exactSolitaryBoundaryPressure(bed, walls, 0.3, 2.5, spacing)
    # Far from the wave, still water: hydrostatic from still water level down to the particle
    inner = np.array([[10.0 - 0.5 * spacing, bed.elevation(10.0 - 0.5 * spacing) + 0.5 * spacing]])
    _, innerPressure, _ = exactSolitaryFlow(inner, 0.3, 1.0, 2.5, 1.0)
    assert pressure[0] == pytest.approx(max(innerPressure[0], 0.0) + g * (inner[0, 1] + 0.5 * spacing), rel = 1e-12)
    # In from the offshore wall at the same height: the water's own pressure, hydrostatic and the wave's
    _, offshore, _ = exactSolitaryFlow(np.array([[0.5 * spacing, 0.5]]), 0.3, 1.0, 2.5, 1.0)
    assert pressure[2] == offshore[0] and g * 0.5 < pressure[2] < g * (0.5 + 0.3)
    assert pressure[3] == 0.0
    assert pressure[1] > g * (1.0 + 1.5 * spacing)


def testTheGrilliStartHoldsTheReferencesParticles():
    from pysurf_amd.scenarios import exactSolitaryOnSlope
    case = exactSolitaryOnSlope(64)
    assert case.fluidCount == 128947
    assert case.smoothingLength == pytest.approx(3.0 / 64.0)
    fluid = case.particles.isFluid
    # Every fluid particle's density carries the wave's pressure, the crest's water at rest density or above
    assert case.particles.densities[fluid].min() >= case.referenceDensity


def flatField():
    from pysurf_amd.bathymetry import HeightField
    return HeightField.fromFunction(lambda x, y: np.zeros_like(x), (0.0, 8.0, -0.25, 0.25), 0.125, stillWaterLevel = 1.0)


@pytest.mark.parametrize('bedFitted', [False, True])
def testOverAFlatBedThe3DWaveIsThe2DWaveLaidAcross(bedFitted):
    from pysurf_amd.exact_wave import exactSolitaryWaveFluid3D
    spacing = 1.0 / 16.0
    positions, velocities, pressures = exactSolitaryWaveFluid3D(flatField(), 0.3, 1.0, 3.0, spacing, bedFitted = bedFitted)
    plane, planeVelocities, planePressures = exactSolitaryWaveFluid(Bathymetry([0.0, 8.0], [0.0, 0.0], 1.0), 0.3, 3.0,
                                                                    spacing)
    rows = np.unique(positions[:, 1])
    assert rows == pytest.approx(np.arange(-0.25 + 0.5 * spacing, 0.25, spacing))
    order = lambda points: np.lexsort(points.T[::-1])
    expected = order(plane)
    for y in rows:
        row = np.flatnonzero(positions[:, 1] == y)
        mine = row[order(positions[row][:, [0, 2]])]
        assert positions[mine][:, [0, 2]] == pytest.approx(plane[expected], abs = 1e-12)
        assert velocities[mine][:, [0, 2]] == pytest.approx(planeVelocities[expected], abs = 1e-12)
        assert np.all(velocities[mine][:, 1] == 0.0)
        assert pressures[mine] == pytest.approx(planePressures[expected], abs = 1e-12)


def testAJittered3DWaveCarriesTheFlowWhereItLies():
    from pysurf_amd.exact_wave import exactSolitaryWaveFluid3D
    spacing = 1.0 / 16.0
    field = flatField()
    laid, _, _ = exactSolitaryWaveFluid3D(field, 0.3, 1.0, 3.0, spacing, bedFitted = True)
    positions, velocities, pressures = exactSolitaryWaveFluid3D(field, 0.3, 1.0, 3.0, spacing, bedFitted = True,
                                                                jitter = 0.15, seed = 2)
    flow, pressure, surface = exactSolitaryFlow(positions[:, [0, 2]], 0.3, 1.0, 3.0, 1.0)
    assert len(positions) == len(laid)
    assert np.all(positions[:, 2] >= 0.5 * spacing - 1e-12) and np.all(positions[:, 2] < 1.0 + surface)
    assert np.all(np.abs(positions[:, 1]) <= 0.25)
    assert velocities[:, [0, 2]] == pytest.approx(flow, abs = 1e-12) and np.all(velocities[:, 1] == 0.0)
    assert pressures == pytest.approx(np.maximum(pressure, 0.0), abs = 1e-12)


def testThe3DWallsFollowThe2DRuleInFromTheSides():
    from pysurf_amd.exact_wave import exactSolitaryBoundaryPressure3D
    spacing = 1.0 / 16.0
    walls = np.array([[3.0, 0.6, 0.4], [3.0, -0.1, -1.5 * spacing], [-0.5 * spacing, 0.0, 0.5]])
    pressure3D = exactSolitaryBoundaryPressure3D(flatField(), walls, 0.3, 1.0, 3.0, spacing)
    pressure2D = exactSolitaryBoundaryPressure(Bathymetry([0.0, 8.0], [0.0, 0.0], 1.0), walls[:, [0, 2]], 0.3, 3.0,
                                               spacing)
    assert pressure3D == pytest.approx(pressure2D, rel = 1e-12)


def testTheExactRidgeCarriesItsPressureIntoTheSplitZone():
    from pysurf_amd.scenarios import ridge3D
    case = ridge3D(particlesPerDepth = 8, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5, ridgeWidth = 0.6,
                   zoneHalfWidth = 0.4, lengthRatio = 0.5, waveHeight = 0.3, start = 'exact', jitter = 0.0)
    whole = ridge3D(particlesPerDepth = 8, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5, ridgeWidth = 0.6,
                    zoneHalfWidth = None, waveHeight = 0.3, start = 'exact', jitter = 0.0)
    p, q = case.particles, whole.particles
    count = q.nParticles
    mothers = np.flatnonzero(q.isFluid & (np.abs(q.positions[:, 0] - 2.5) < 0.4))
    assert p.nParticles == count + 7 * len(mothers) and len(mothers) > 0
    # The first daughter keeps its mother's slot and density; the other seven follow, mother by mother
    assert np.array_equal(p.densities[:count], q.densities)
    assert np.array_equal(p.densities[count:], np.repeat(q.densities[mothers], 7))
    with pytest.raises(ValueError):
        ridge3D(particlesPerDepth = 8, length = 6.0, width = 1.5, start = 'serre')
