'''

The breaking read-out measured on the device (sph_surface_profile, sph_face_cells) against the
tracker's own measures of the downloaded positions: the free-surface profile, the face window's
marked cells and the waterline must agree bit for bit on 2D runs and a 3D ridge run as they step,
also where the band limit is cut so the host fallback decides; and a run tracked from device
measures must record the history it records from its positions.

'''

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def assertSameMeasures(device: dict, host: dict) -> None:
    assert device.keys() == host.keys()
    for run in host:
        if host[run] is None:
            assert device[run] is None, run
            continue
        centersD, elevationD, occupiedD, startD, tipD = device[run]
        centersH, elevationH, occupiedH, startH, tipH = host[run]
        assert np.array_equal(centersD, centersH), run
        assert np.array_equal(elevationD, elevationH, equal_nan = True), run
        assert startD == startH, run
        assert tipD == tipH or (np.isnan(tipD) and np.isnan(tipH)), (run, tipD, tipH)
        # The host keeps bands up to its highest marked one, the device a fixed number: empty bands
        # above the water carry no face
        assert occupiedD.shape[1] == occupiedH.shape[1], run
        rows = max(occupiedD.shape[0], occupiedH.shape[0])
        pad = lambda o: np.vstack([o, np.zeros((rows - o.shape[0], o.shape[1]), dtype = bool)])
        assert np.array_equal(pad(occupiedD), pad(occupiedH)), run
        assert occupiedH.any(), run


def cases(dimensions: int) -> list:
    from pysurf_amd.scenarios import ridge3D, solitaryOnSlope
    if dimensions == 2:
        return [solitaryOnSlope(particlesPerDepth = 8, seed = 1), solitaryOnSlope(particlesPerDepth = 10, seed = 2)]
    return [ridge3D(particlesPerDepth = 8, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5, ridgeWidth = 0.6,
                    zoneHalfWidth = 0.4, lengthRatio = 0.5)]


@pytest.mark.parametrize('dimensions', [2, 3])
def testDeviceMeasuresAreTheTrackersMeasures(dimensions):
    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.readout import DeviceReadout, hostMeasures
    from pysurf_amd.serving import stripOf, trackerOf
    runs = cases(dimensions)
    batch = BatchEngine(runs, 'float32', lanes = 2, listCapacity = 64 if dimensions == 2 else 160)
    batch.setClocks()
    entries = [(r, trackerOf(case), *stripOf(case)) for r, case in enumerate(runs)]
    readout = DeviceReadout()
    for reading in range(4):
        if reading:
            batch.advance(40)
        host = hostMeasures(batch, entries)
        device = readout.measure(batch, entries)
        print(f'{dimensions}D reading {reading}: tips {[host[r][4] for r in host]}', flush = True)
        assertSameMeasures(device, host)
        assert readout.fallbacks == 0
        # Bands cut below the water: every run overflows and the host measures it
        cut = readout.measure(batch, entries, bandMargin = -40)
        assert readout.fallbacks == len(entries)
        readout.fallbacks = 0
        assertSameMeasures(cut, host)


def testARunTrackedOnTheDeviceRecordsTheHistoryOfItsPositions():
    from pysurf_amd.scenarios import solitaryOnSlope
    from pysurf_amd.serving import runToBreaking
    case = solitaryOnSlope(particlesPerDepth = 8, seed = 1)
    device = runToBreaking([case], 'float32', endTime = 0.3, listCapacity = 64, measure = 'device')[0]
    host = runToBreaking([case], 'float32', endTime = 0.3, listCapacity = 64, measure = 'host')[0]
    print(f'{len(device.times)} readings', flush = True)
    assert len(device.times) > 10
    for name in ('times', 'crestX', 'crestHeight', 'lean', 'waterline'):
        assert np.array_equal(np.asarray(getattr(device, name)), np.asarray(getattr(host, name)), equal_nan = True), name
    assert device.maxRunup == host.maxRunup or (np.isnan(device.maxRunup) and np.isnan(host.maxRunup))
