'''

Grilli's 1:35 case from the exact solitary wave (pysurf_amd.scenarios.exactSolitaryOnSlope, the
reference's BreakingWave 'deltaSph' start: bed-fitted columns, Wendland kernel at h = 3 spacings),
stepped on the GPU in float32 pairs at 64 particles per depth and read by the smoothed-surface
tracker until the jet's touchdown fixes the breaking point. It must break where the reference's own
GPU run recorded it (tests/golden/grilli_breaking_64.json, from tests/golden/make_grilli_record.py),
and at 128 particles per depth with the renormalized 'antuono' diffusion, where the reference's run
with it recorded it (tests/golden/grilli_breaking_128_antuono.json):
the time, crest height and depths within 1% (the recorded run's own scatter under rounding-level
changes is a few tenths of a percent on this measure), and within the reference's criteria against
Grilli et al. (1997): Hb/h0 within 10%, hb/h0 within 15%.

'''

import json
import os
import time

import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
# (particles per depth, density diffusion, the recorded run)
RECORDS = [(64, 'hydrostatic', 'grilli_breaking_64.json'), (128, 'antuono', 'grilli_breaking_128_antuono.json')]


@pytest.mark.timeout(900)
@pytest.mark.parametrize('ppd, diffusion, record', RECORDS)
def testTheGrilliCaseBreaksWhereTheReferenceRecordedIt(ppd, diffusion, record):
    from pysurf_amd.scenarios import exactSolitaryOnSlope
    from pysurf_amd.serving import Fleet, Member
    golden = json.load(open(os.path.join(GOLDEN, record)))
    recorded, grilli = golden['recorded'], golden['grilli1997']
    print(f'Grilli 1:35 at {ppd} ppd, {diffusion} diffusion', flush = True)
    case = exactSolitaryOnSlope(ppd)
    case.diffusion = diffusion
    assert case.fluidCount == recorded['fluidParticles']
    fleet = Fleet(2, 'float32', lanes = 4, listCapacity = 160, chunk = 8)
    member = Member(key = 0, case = case, endTime = 12.0)
    fleet.rebuild([], [member])
    fleet.launch()
    started, readings = time.perf_counter(), 0
    while not member.done:
        reading = fleet.collect()
        if not member.done:
            fleet.launch()
        fleet.track(reading, time.perf_counter())
        readings += 1
        if readings % 500 == 0:
            print(f'{readings} readings: t = {member.time:.3f} s, {member.steps} steps, '
                  f'{time.perf_counter() - started:.1f} s', flush = True)
    tracker = member.tracker
    point = tracker.breakingPoint
    assert tracker.breakingFinal and point is not None
    print(f'breaking at {point.time:.4f} s: Hb/h0 {point.crestHeight:.4f}, hb/h0 {point.depth:.4f}, face depth '
          f'{point.faceDepth:.4f}; touchdown {tracker.touchdownTime:.4f} s; {member.steps} steps in '
          f'{time.perf_counter() - started:.1f} s', flush = True)
    assert point.time == pytest.approx(recorded['breakingTime'], rel = 0.01)
    assert tracker.touchdownTime == pytest.approx(recorded['touchdownTime'], rel = 0.01)
    assert point.crestHeight == pytest.approx(recorded['breakingHeightOverDepth'], rel = 0.01)
    assert point.depth == pytest.approx(recorded['breakingDepthOverDepth'], rel = 0.01)
    assert point.faceDepth == pytest.approx(recorded['breakingFaceDepthOverDepth'], rel = 0.01)
    assert abs(point.crestHeight / grilli['breakingHeightOverDepth'] - 1.0) <= 0.10
    assert abs(point.depth / grilli['breakingDepthOverDepth'] - 1.0) <= 0.15


def testAnExactWaveKeepsItsFormAndSpeedIn3D():
    # The exact wave of H/d = 0.3 over the flat part of a ridge tank, stepped 0.3 s on the GPU:
    # along the tank's centre strip its highest column stays within half a spacing (the profile's
    # resolution) and 3% of H, and the centroid of its upper half travels at the exact celerity
    # within 3%
    import numpy as np
    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.exact_wave import exactSolitaryWaveSpeed
    from pysurf_amd.readout import hostMeasures
    from pysurf_amd.scenarios import ridge3D
    from pysurf_amd.serving import Member, readingEntry
    case = ridge3D(particlesPerDepth = 16, length = 8.0, width = 1.0, toeX = 5.5, crestX = 2.5, ridgeWidth = 0.6,
                   zoneHalfWidth = None, waveHeight = 0.3, start = 'exact', jitter = 0.0)
    batch = BatchEngine([case], 'float32', lanes = 4, listCapacity = 256)
    batch.setClocks(untilTime = 0.3)
    entry = readingEntry(0, Member(key = 0, case = case, endTime = 1.0))

    def wave() -> tuple:
        measures = hostMeasures(batch, [entry])[0]
        excess = np.nan_to_num(measures['elevation'] - 1.0, nan = 0.0)
        # The wave's upper half, symmetric about its crest, clear of the disturbance the offshore
        # wall sends after it
        upper = np.where(excess > 0.5 * excess.max(), excess, 0.0)
        return float(excess.max()), float(np.sum(measures['centers'] * upper) / np.sum(upper))

    _, startCentroid = wave()
    while batch.clocks()['time'][0] < 0.3:
        batch.advance(64)
    elapsed = float(batch.clocks()['time'][0])
    height, centroid = wave()
    travelled, expected = centroid - startCentroid, exactSolitaryWaveSpeed(0.3, 1.0) * elapsed
    print(f'after {elapsed:.3f} s: crest {height:.4f} high, centroid moved {travelled:.4f}, {expected:.4f} expected',
          flush = True)
    assert abs(height - 0.3) <= 0.5 * case.spacing + 0.03 * 0.3
    assert travelled == pytest.approx(expected, rel = 0.03)
