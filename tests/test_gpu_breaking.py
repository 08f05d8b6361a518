'''

Breaking on the device against the float64 restatement, at 16 and 64 particles per depth, and the
request loop's promise that a run served in a mix is the run served alone.

The 16-particles-per-depth solitary wave on a 1:15 beach is stepped on the GPU to its breaking point
with float32 pairs and with the plain float64 kernel, read by the same tracker every 0.01 sqrt(d/g)
as the oracle run in tests/golden/breaking_c1_16ppd.json (tests/golden/make_breaking.py): the
breaking time and the crest's position at breaking must agree within 1%, as means over an ensemble
of starts moved by rounding-level amounts, because single trajectories scatter by about 2% under
rounding-level changes.

'''

import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'breaking_c1_16ppd.json')


def perturbed(case, seed: int, size: float = 1e-9):
    '''The case with its fluid moved by up to `size` spacings per axis (seed 0: as built): a rounding-level perturbation.'''
    if seed == 0:
        return case
    particles = case.particles.copy()
    fluid = particles.isFluid
    rng = np.random.default_rng(seed)
    particles.positions[fluid] += rng.uniform(-size, size, (int(fluid.sum()), case.dimensions)) * case.spacing
    return type(case)(**{**case.__dict__, 'particles': particles})


def ensembleBreaking(case, members: int, precision: str) -> np.ndarray:
    '''(time, crest x) of the breaking point of each member of the perturbed ensemble, stepped as one batch.'''
    from pysurf_amd.serving import runToBreaking
    trackers = runToBreaking([perturbed(case, seed) for seed in range(members)], precision, endTime = 10.0)
    assert all(tracker.breakingPoint is not None for tracker in trackers)
    return np.array([(t.breakingPoint.time, t.breakingPoint.crestX) for t in trackers])


@pytest.mark.timeout(300)
def testBreakingTimeAndLocationFollowTheOracle():
    # A single trajectory's breaking time read in cells moves by about 2% under rounding-level
    # changes, in float64 as in float32. The float32 engine and the plain float64 kernel therefore
    # step the same ensemble of 8 starts moved by up to 1e-9 spacings, and their means must agree
    # within 1% in time and place; the oracle's own run must lie within the float64 ensemble's reach.
    from pysurf_amd.scenarios import solitaryOnSlope
    expected = json.load(open(GOLDEN))
    case = solitaryOnSlope(particlesPerDepth = 16)
    samples = {precision: ensembleBreaking(case, 8, precision) for precision in ('float64', 'float32')}
    for precision, values in samples.items():
        print(f'{precision}: times {values[:, 0].round(4).tolist()}, crest x {values[:, 1].round(3).tolist()}', flush = True)
    double, single = samples['float64'].mean(axis = 0), samples['float32'].mean(axis = 0)
    print(f'means: float64 {double}, float32 {single}, oracle ({expected["time"]}, {expected["crestX"]})', flush = True)
    assert single[0] == pytest.approx(double[0], rel = 0.01)
    assert single[1] == pytest.approx(double[1], rel = 0.01)
    spread = samples['float64'][:, 0].max() - samples['float64'][:, 0].min()
    assert abs(expected['time'] - double[0]) <= max(spread, 0.01 * double[0])


def testARunServedInAMixIsTheRunServedAlone():
    # Two 2D runs and a 3D ridge run arrive together, one 2D run finishes early and leaves, so the
    # fleets rebuild mid-run; every run's tracker history must match the run's own, alone, bit for bit
    from pysurf_amd.scenarios import ridge3D, solitaryOnSlope
    from pysurf_amd.serving import Request, Server, runToBreaking
    cases = [solitaryOnSlope(particlesPerDepth = 8, seed = 1), solitaryOnSlope(particlesPerDepth = 10, seed = 2),
             ridge3D(particlesPerDepth = 8, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5, ridgeWidth = 0.6,
                     zoneHalfWidth = 0.4, lengthRatio = 0.5)]
    ends = [0.15, 0.3, 0.2]
    requests = [Request(key = k, arrival = 0.0, case = case, endTime = end) for k, (case, end) in enumerate(zip(cases, ends))]
    server = Server(budgetBytes = 1e12, listCapacity3D = 160)
    served = {record.key: record for record in server.serve(requests)}
    assert sorted(served) == [0, 1, 2]
    for k, (case, end) in enumerate(zip(cases, ends)):
        alone = runToBreaking([case], 'float32', endTime = end, lanes = server.lanes,
                              listCapacity = 64 if case.dimensions == 2 else 160)[0]
        mixed = served[k].tracker
        print(f'run {k}: {len(alone.times)} readings alone, {len(mixed.times)} in the mix', flush = True)
        assert served[k].steps > 0 and len(alone.times) > 10
        for name in ('times', 'crestX', 'crestHeight', 'lean', 'faceSlope', 'waterline'):
            assert np.array_equal(np.asarray(getattr(alone, name)), np.asarray(getattr(mixed, name)), equal_nan = True), \
                (k, name)


GOLDEN64 = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'breaking_c2_64ppd.json')


@pytest.mark.timeout(600)
def testTheSingle64ppdRunBreaksWhereTheOracleBreaks():
    # The single-run configuration (64 particles per depth, H/d and slope drawn from seed 0): the
    # float32-pair engine and the plain float64 kernel each step the perturbed ensemble of 8 starts to
    # its breaking point, and each precision's mean breaking time and crest position must lie within
    # 1% of the float64 oracle's mean over the same members (tests/golden/breaking_c2_64ppd.json,
    # tests/golden/make_breaking_64.py), and within 1% of each other
    from pysurf_amd.scenarios import singleRun
    expected = json.load(open(GOLDEN64))
    members = expected['members']
    assert len(members) >= 3
    oracleMean = np.array([np.mean([m['time'] for m in members]), np.mean([m['crestX'] for m in members])])
    case = singleRun(seed = 0, particlesPerDepth = 64)
    samples = {precision: ensembleBreaking(case, 8, precision) for precision in ('float64', 'float32')}
    for precision, values in samples.items():
        print(f'{precision}: times {values[:, 0].round(4).tolist()}, crest x {values[:, 1].round(3).tolist()}', flush = True)
    print(f'oracle ({len(members)} members): times {[round(m["time"], 4) for m in members]}, '
          f'crest x {[round(m["crestX"], 3) for m in members]}', flush = True)
    double, single = samples['float64'].mean(axis = 0), samples['float32'].mean(axis = 0)
    print(f'means: float64 {double}, float32 {single}, oracle {oracleMean}', flush = True)
    for mean in (double, single):
        assert mean[0] == pytest.approx(oracleMean[0], rel = 0.01)
        assert mean[1] == pytest.approx(oracleMean[1], rel = 0.01)
    assert single[0] == pytest.approx(double[0], rel = 0.01)
    assert single[1] == pytest.approx(double[1], rel = 0.01)


CREST = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'crest_c1_16ppd.json')


@pytest.mark.timeout(300)
def testTheCrestTravelsAsInTheOracle():
    # The crest's speed, fitted to its track over 0.5-3 s (the flat bed, the toe at 1.1 s and the
    # slope), and its mean height there, in float32 pairs and the plain float64 kernel, against the
    # float64 restatement's track (tests/golden/crest_c1_16ppd.json, tests/golden/make_crest.py):
    # speed within 1% over the whole window and 2% over its last second, height within 2%
    from pysurf_amd.scenarios import solitaryOnSlope
    from pysurf_amd.serving import runToBreaking
    expected = json.load(open(CREST))

    def measures(times, crestX, crestHeight) -> list:
        t, x, h = np.asarray(times), np.asarray(crestX), np.asarray(crestHeight)
        out = []
        for low, high in ((0.5, 3.0), (2.0, 3.0)):
            window = (t >= low) & (t <= high)
            out.append(float(np.polyfit(t[window], x[window], 1)[0]))
        window = (t >= 0.5) & (t <= 3.0)
        out.append(float(h[window].mean()))
        return out

    oracleSpeed, oracleLate, oracleHeight = measures(expected['time'], expected['crestX'], expected['crestHeight'])
    for precision in ('float64', 'float32'):
        tracker = runToBreaking([solitaryOnSlope(particlesPerDepth = 16)], precision, endTime = 3.0)[0]
        speed, late, height = measures(tracker.times, tracker.crestX, tracker.crestHeight)
        print(f'{precision}: crest speed {speed:.4f} (oracle {oracleSpeed:.4f}), last second {late:.4f} '
              f'({oracleLate:.4f}), mean height {height:.4f} ({oracleHeight:.4f})', flush = True)
        assert speed == pytest.approx(oracleSpeed, rel = 0.01)
        assert late == pytest.approx(oracleLate, rel = 0.02)
        assert height == pytest.approx(oracleHeight, rel = 0.02)
