'''

Splitting on the device (sph_split_count, sph_split_apply through BatchEngine.split) against the
host's splitParticles (pysurf_amd/refinement.py, after sphRefinement.py:135-186) on the same state:
every particle's position to rounding (the daughters' offsets come from a root of the mother's
float32 mass on the device, of her float64 mass on the host), and velocity, density, pressure,
mass, length and kind exactly; the run's particles in the host's order (its own, then the other
daughters mother by mother). A batch split on the device then steps as the host-split runs do.

'''

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def hostSplit(case, zone, lengthRatio):
    from pysurf_amd.refinement import splitParticles
    particles = case.particles.copy()
    x = particles.positions[:, 0]
    mask = particles.isFluid & (x >= zone[0]) & (x < zone[1])
    splitParticles(particles, mask, case.referenceDensity, case.smoothingLength, lengthRatio)
    return type(case)(**{**case.__dict__, 'particles': particles}), int(mask.sum())


def cases(dimensions):
    from pysurf_amd.scenarios import ridge3D, solitaryOnSlope
    if dimensions == 2:
        return ([solitaryOnSlope(particlesPerDepth = 8, seed = 1), solitaryOnSlope(particlesPerDepth = 10, seed = 2)],
                [(3.5, 4.5), None], 1.0)
    return ([ridge3D(particlesPerDepth = 8, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5, ridgeWidth = 0.6,
                     zoneHalfWidth = None)], [(2.1, 2.9)], 0.5)


@pytest.mark.parametrize('dimensions', [2, 3])
def testTheDeviceSplitIsTheHostSplit(dimensions):
    from pysurf_amd.batch import BatchEngine
    runs, zones, ratio = cases(dimensions)
    capacity = 64 if dimensions == 2 else 256
    batch = BatchEngine(runs, 'float32', lanes = 2, listCapacity = capacity)
    batch.setClocks(untilTime = 1.0)
    before = [batch.offsets[r + 1] - batch.offsets[r] for r in range(len(runs))]
    chosen = batch.split(zones, lengthRatio = ratio)
    expected = []
    for r, (case, zone) in enumerate(zip(runs, zones)):
        if zone is None:
            expected.append((case, 0))
        else:
            expected.append(hostSplit(case, zone, ratio))
    print(f'{dimensions}D: split {chosen.tolist()} of {before}', flush = True)
    assert chosen.tolist() == [count for _, count in expected] and chosen.sum() > 0
    assert batch.count == sum(case.count for case, _ in expected)
    for r, (case, _) in enumerate(expected):
        state, host = batch.state(r), case.particles
        assert state.nParticles == host.nParticles
        assert np.abs(state.positions - host.positions).max() < 1e-9, r
        assert np.array_equal(state.velocities, host.velocities), r
        assert np.array_equal(state.densities, host.densities), r
        assert np.array_equal(state.isFluid, host.isFluid), r
        if chosen[r] == 0:
            # A run that did not split keeps its case's masses and lengths
            assert np.array_equal(state.masses, host.masses) and state.lengths is host.lengths is None, r
            continue
        # The device holds masses and lengths in the pair precision, float32 here
        assert np.array_equal(state.masses, host.masses.astype(np.float32).astype(np.float64)), r
        lengths = host.lengths if host.lengths is not None else np.full(host.nParticles, case.smoothingLength)
        assert np.array_equal(state.lengths, lengths.astype(np.float32).astype(np.float64)), r
    # The clocks carried over, and the split batch steps as the host-split runs do
    split = BatchEngine([case for case, _ in expected], 'float32', lanes = 2, listCapacity = capacity)
    split.setClocks(untilTime = 1.0)
    batch.advance(20)
    split.advance(20)
    assert np.array_equal(batch.clocks()['steps'], split.clocks()['steps'])
    for r in range(len(runs)):
        mine, theirs = batch.state(r), split.state(r)
        assert np.abs(mine.positions - theirs.positions).max() < 1e-6, r
        assert np.abs(mine.densities - theirs.densities).max() < 1e-3 * 1000.0, r


def testASortedBatchRefusesToSplit():
    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.scenarios import solitaryOnSlope
    batch = BatchEngine([solitaryOnSlope(particlesPerDepth = 8, seed = 1)], 'float32', lanes = 2, listCapacity = 64)
    batch.sortParticles()
    with pytest.raises(ValueError):
        batch.split([(3.5, 4.5)])
