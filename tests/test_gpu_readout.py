'''

The breaking read-out measured on the device (sph_surface_profile, sph_face_cells,
sph_fluid_fraction) against the tracker's own measures of the downloaded positions: the free-surface
profile, the face window's marked cells and the waterline must agree bit for bit, and the smoothed fluid fraction a 3D run's
tracker reads to 1e-12, on 2D runs and a 3D ridge run as they step, also where the band limit is
cut so the host fallback decides; the fraction read twice is the same bits; and a run tracked from
device measures must record the history it records from its positions.

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
        d, h = device[run], host[run]
        assert d.keys() == h.keys(), (run, d.keys(), h.keys())
        assert np.array_equal(d['centers'], h['centers']), run
        assert np.array_equal(d['elevation'], h['elevation'], equal_nan = True), run
        assert d['tip'] == h['tip'] or (np.isnan(d['tip']) and np.isnan(h['tip'])), (run, d['tip'], h['tip'])
        if 'occupied' in h:
            assert d['windowStart'] == h['windowStart'], run
            # The host keeps bands up to its highest marked one, the device a fixed number: empty
            # bands above the water carry no face
            occupiedD, occupiedH = d['occupied'], h['occupied']
            assert occupiedD.shape[1] == occupiedH.shape[1], run
            rows = max(occupiedD.shape[0], occupiedH.shape[0])
            pad = lambda o: np.vstack([o, np.zeros((rows - o.shape[0], o.shape[1]), dtype = bool)])
            assert np.array_equal(pad(occupiedD), pad(occupiedH)), run
            assert occupiedH.any(), run
        if 'grid' in h:
            assert (d['grid'] is None) == (h['grid'] is None), run
            if h['grid'] is not None:
                (xD, yD, phiD), (xH, yH, phiH) = d['grid'], h['grid']
                assert np.array_equal(xD, xH) and np.array_equal(yD, yH), run
                # The device's root and powers may round apart from numpy's
                assert np.abs(phiD - phiH).max() <= 1e-12, (run, np.abs(phiD - phiH).max())
                assert phiH.max() > 0.9, run


def cases(dimensions: int) -> list:
    from pysurf_amd.scenarios import ridge3D, solitaryOnSlope
    if dimensions == 2:
        return [solitaryOnSlope(particlesPerDepth = 8, seed = 1), solitaryOnSlope(particlesPerDepth = 10, seed = 2)]
    # 16 particles per depth: at 8 the window's tenth of a depth is less than a spacing, and there is none
    return [ridge3D(particlesPerDepth = 16, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5, ridgeWidth = 0.6,
                    zoneHalfWidth = 0.4, lengthRatio = 0.5)]


@pytest.mark.parametrize('dimensions', [2, 3])
def testDeviceMeasuresAreTheTrackersMeasures(dimensions):
    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.readout import DeviceReadout, hostMeasures
    from pysurf_amd.serving import Member, readingEntry
    runs = cases(dimensions)
    batch = BatchEngine(runs, 'float32', lanes = 2, listCapacity = 64 if dimensions == 2 else 768)
    batch.setClocks()
    entries = [readingEntry(r, Member(key = r, case = case, endTime = 1.0)) for r, case in enumerate(runs)]
    readout = DeviceReadout()
    grids = 0
    for reading in range(4):
        if reading:
            batch.advance(40)
        host = hostMeasures(batch, entries)
        device = readout.measure(batch, entries)
        print(f'{dimensions}D reading {reading}: tips {[host[r]["tip"] for r in host]}', flush = True)
        assertSameMeasures(device, host)
        assert readout.fallbacks == 0
        # The same state read again gives the same bits
        again = readout.measure(batch, entries)
        for run in device:
            if 'grid' in device[run] and device[run]['grid'] is not None:
                assert np.array_equal(again[run]['grid'][2], device[run]['grid'][2])
                grids += 1
        if dimensions == 2:
            # Bands cut below the water: every run overflows and the host measures it
            cut = readout.measure(batch, entries, bandMargin = -40)
            assert readout.fallbacks == len(entries)
            readout.fallbacks = 0
            assertSameMeasures(cut, host)
    assert grids == (4 if dimensions == 3 else 0)


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


def testA3DRunReadOnTheDeviceRecordsTheHistoryOfItsPositions():
    # The smoothed surface read from the device's fraction and from the host's agree to rounding
    from pysurf_amd.serving import runToBreaking
    case = cases(3)[0]
    device = runToBreaking([case], 'float32', endTime = 0.2, listCapacity = 768, measure = 'device')[0]
    host = runToBreaking([case], 'float32', endTime = 0.2, listCapacity = 768, measure = 'host')[0]
    slopes = np.asarray(device.faceSlope)
    print(f'{len(device.times)} readings, {int(np.isfinite(slopes).sum())} face slopes', flush = True)
    assert device.faceMeasure == 'contour' and len(device.times) > 5 and np.isfinite(slopes).sum() > 5
    for name in ('times', 'lean', 'waterline'):
        assert np.array_equal(np.asarray(getattr(device, name)), np.asarray(getattr(host, name)), equal_nan = True), name
    for name in ('crestX', 'crestHeight', 'faceSlope'):
        assert np.allclose(getattr(device, name), getattr(host, name), rtol = 0.0, atol = 1e-9, equal_nan = True), name
