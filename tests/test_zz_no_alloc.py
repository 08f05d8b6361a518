'''

sph_state_sums, the global sums that allocate nothing (their buffers live in one file-scope device
array), against sph_global_sums of the same runs and state: bit for bit, on two calls in a row, the
second over a smaller state (fewer runs, fewer particles), so it also shows that what the first
call left in the shared array does not leak into the second. Limits are refused with a message.

'''

import ctypes
import sys

import numpy as np
import pytest

from pysurf_amd import _native

pytestmark = pytest.mark.gpu


def contextFree(batch, runs=None) -> np.ndarray:
    '''sph_state_sums over the batch's bound state and its runs (or the first `runs` of them).'''
    engine = batch.engine
    every = list(engine._runs)
    chosen = every if runs is None else every[:runs]
    array = (_native.Run * len(chosen))(*chosen)
    state = engine.stateOf(engine.tensors, engine.capacity)
    sums = np.full((len(chosen), len(_native.globalSumNames)), np.nan)
    status = engine.library.sph_state_sums(engine.dimensions, 0 if engine.precision == 'float32' else 1, array,
                                           len(chosen), ctypes.byref(state), sums.ctypes.data_as(ctypes.c_void_p),
                                           ctypes.c_void_p(engine.stream.cuda_stream))
    if status != _native.SPH_OK:
        raise ValueError(engine.library.sph_static_error().decode())
    return sums


def stepped(cases, steps, precision):
    from pysurf_amd.batch import BatchEngine
    batch = BatchEngine(cases, precision)
    batch.setClocks(maxSteps = steps)
    batch.advance(steps)
    batch.synchronize()
    return batch


@pytest.mark.parametrize('precision', ['float32', 'float64'])
def testContextFreeSumsEqualTheContextSumsTwiceInARow(precision):
    from pysurf_amd.scenarios import solitaryOnSlope
    large = [solitaryOnSlope(particlesPerDepth = ppd, seed = seed) for ppd, seed in ((16, 1), (12, 2), (20, 3))]
    print('no-alloc sums: three 2D runs,', [case.count for case in large], precision, flush = True)
    first = stepped(large, 12, precision)
    expected = first.globalSums()
    got = contextFree(first)
    assert np.array_equal(got, expected), np.abs(got - expected).max()
    # a smaller state: one run of fewer particles, in a batch of its own
    small = [solitaryOnSlope(particlesPerDepth = 6, seed = 4)]
    print('no-alloc sums: one 2D run,', small[0].count, precision, flush = True)
    second = stepped(small, 7, precision)
    expected = second.globalSums()
    got = contextFree(second)
    assert got.shape == (1, 7)
    assert np.array_equal(got, expected), np.abs(got - expected).max()
    # and a prefix of the first batch's runs, over its still-bound state
    got = contextFree(first, runs = 1)
    assert np.array_equal(got, first.globalSums()[:1])


def testContextFreeSumsIn3D():
    from pysurf_amd.scenarios import ridge3D
    case = ridge3D(particlesPerDepth = 8, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5, ridgeWidth = 0.6,
                   zoneHalfWidth = 0.4, lengthRatio = 0.5, seed = 0)
    print('no-alloc sums: one 3D run,', case.count, flush = True)
    batch = stepped([case], 5, 'float32')
    assert np.array_equal(contextFree(batch), batch.globalSums())


def testContextFreeSumsRefuseWhatTheirArrayCannotHold():
    from pysurf_amd.scenarios import solitaryOnSlope
    batch = stepped([solitaryOnSlope(particlesPerDepth = 6, seed = 5)], 1, 'float32')
    engine = batch.engine
    run = list(engine._runs)[0]
    many = (_native.Run * 1025)(*([run] * 1025))
    state = engine.stateOf(engine.tensors, engine.capacity)
    sums = np.zeros((1025, 7))
    status = engine.library.sph_state_sums(2, 0, many, 1025, ctypes.byref(state), sums.ctypes.data_as(ctypes.c_void_p),
                                           ctypes.c_void_p(engine.stream.cuda_stream))
    assert status == _native.SPH_ERR_VALUE
    assert b'1024 runs' in engine.library.sph_static_error()
    outside = _native.Run.from_buffer_copy(run)
    outside.count = engine.capacity + 1
    status = engine.library.sph_state_sums(2, 0, ctypes.byref(outside), 1, ctypes.byref(state),
                                           sums.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(engine.stream.cuda_stream))
    assert status == _native.SPH_ERR_VALUE
    assert b'outside' in engine.library.sph_static_error()
    # the call after a refusal still works
    assert np.array_equal(contextFree(batch), batch.globalSums())


if __name__ == '__main__':
    sys.exit(pytest.main([__file__, '-x', '-v', '-p', 'no:cacheprovider']))
