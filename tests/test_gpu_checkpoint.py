'''

Resuming from a checkpoint on the device: a batch saved midway (after a device sort, so with a
permuted state) and restored into a fresh engine built on the same runs steps on exactly as the
uninterrupted batch does, bit for bit, clocks included; a checkpoint of other runs is refused.

'''

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def cases():
    from pysurf_amd.scenarios import solitaryOnSlope
    return [solitaryOnSlope(particlesPerDepth = 6, seed = 2, waveHeight = 0.3), solitaryOnSlope(particlesPerDepth = 8, seed = 3)]


@pytest.mark.parametrize('precision', ['float32', 'float64'])
def testAResumedBatchStepsAsTheUninterruptedOne(tmp_path, precision):
    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.checkpoint import restore, save
    print(f'checkpoint resume {precision}', flush = True)
    path = str(tmp_path / 'batch.ckpt')
    first = BatchEngine(cases(), precision)
    first.setClocks(untilTime = 5.0)
    first.advance(20, graph = False)
    first.sortParticles()
    first.advance(10)
    save(first, path, extra = {'tracker': [0.25, 0.5]})
    first.advance(25)
    first.synchronize()
    second = BatchEngine(cases(), precision)
    extra = restore(second, path)
    assert extra == {'tracker': [0.25, 0.5]}
    second.advance(25)
    second.synchronize()
    a, b = first.clocks(), second.clocks()
    assert np.array_equal(a['time'], b['time']) and np.array_equal(a['steps'], b['steps'])
    assert list(a['steps']) == [55, 55]
    for index in range(2):
        x, y = first.state(index), second.state(index)
        for name in ('positions', 'velocities', 'densities', 'pressures', 'accelerations'):
            assert np.array_equal(getattr(x, name), getattr(y, name)), (index, name)


def testACheckpointOfOtherRunsIsRefused(tmp_path):
    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.checkpoint import restore, save
    path = str(tmp_path / 'batch.ckpt')
    save(BatchEngine(cases(), 'float32'), path)
    with pytest.raises(ValueError, match = 'other runs'):
        restore(BatchEngine(cases()[:1], 'float32'), path)
    with pytest.raises(ValueError, match = 'other runs'):
        restore(BatchEngine(cases(), 'float64'), path)


def testACheckpointForAnotherBedOrOtherPhysicsIsRefused(tmp_path):
    # The same particle counts over another bed, or with another sound speed: the run digests differ
    import dataclasses

    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.checkpoint import restore, save
    from pysurf_amd.scenarios import solitaryOnSlope
    path = str(tmp_path / 'batch.ckpt')
    runs = cases()
    save(BatchEngine(runs, 'float32'), path)
    steeper = solitaryOnSlope(particlesPerDepth = 8, seed = 3, slope = 1.0 / 12.0)
    otherBed = [runs[0], dataclasses.replace(runs[1], boundary = steeper.boundary)]
    assert otherBed[1].count == runs[1].count
    with pytest.raises(ValueError, match = 'another bed'):
        restore(BatchEngine(otherBed, 'float32'), path)
    louder = [runs[0], dataclasses.replace(runs[1], soundSpeed = 1.05 * runs[1].soundSpeed)]
    with pytest.raises(ValueError, match = 'other physics'):
        restore(BatchEngine(louder, 'float32'), path)
    # the batch it was written from still restores
    restore(BatchEngine(cases(), 'float32'), path)
