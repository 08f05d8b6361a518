'''

Merging on the device (sph_merge through BatchEngine.merge) against the host's mergeParticles
(pysurf_amd/refinement.py, after sphRefinement.py:212-311) on the same stepped state: the same
groups from the same float64 positions, the same particles left in the same order (the holes among
the boundary's places filled from the end), their positions, velocities and densities to rounding
(the mass-weighted means are summed in another order), masses and kinds exactly, and the same
kinetic energy removed. A merged batch then steps on.

'''

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def splitCase(dimensions):
    from pysurf_amd.refinement import splitParticles
    from pysurf_amd.scenarios import ridge3D, solitaryOnSlope
    if dimensions == 2:
        case, zone, ratio = solitaryOnSlope(particlesPerDepth = 8, seed = 1), (3.0, 5.0), 1.0
    else:
        case, zone, ratio = ridge3D(particlesPerDepth = 8, length = 6.0, width = 1.5, toeX = 2.0, crestX = 2.5,
                                    ridgeWidth = 0.6, zoneHalfWidth = None), (2.0, 3.0), 0.5
    particles = case.particles.copy()
    unsplit = float(particles.masses[particles.isFluid][0])
    x = particles.positions[:, 0]
    splitParticles(particles, particles.isFluid & (x >= zone[0]) & (x < zone[1]), case.referenceDensity,
                   case.smoothingLength, ratio)
    return type(case)(**{**case.__dict__, 'particles': particles}), zone, ratio, unsplit


@pytest.mark.parametrize('dimensions', [2, 3])
def testTheDeviceMergeIsTheHostMerge(dimensions):
    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.refinement import mergeParticles
    case, zone, ratio, unsplit = splitCase(dimensions)
    capacity = 64 if dimensions == 2 else 256
    batch = BatchEngine([case], 'float32', lanes = 2, listCapacity = capacity)
    batch.setClocks(untilTime = 1.0)
    batch.advance(30)
    host = batch.state(0)
    # The device holds masses and lengths in float32: the host merges what the device holds
    host.masses = host.masses.astype(np.float32).astype(np.float64)
    if host.lengths is not None:
        host.lengths = host.lengths.astype(np.float32).astype(np.float64)
    steps = int(batch.clocks()['steps'][0])
    x = host.positions[:, 0]
    expected = mergeParticles(host, (x >= zone[0]) & (x < zone[1]), case.referenceDensity, unsplit,
                              case.smoothingLength, ratio)
    result = batch.merge([zone], unsplit, ratio)[0]
    print(f'{dimensions}D: {result.groups} groups merged on the device, {expected.groups} on the host; energy '
          f'{result.removedEnergy:.6e} against {expected.removedEnergy:.6e}', flush = True)
    assert result.groups == expected.groups > 0
    assert result.removedEnergy == pytest.approx(expected.removedEnergy, rel = 1e-9)
    assert batch.count == host.nParticles and int(batch.clocks()['steps'][0]) == steps
    state = batch.state(0)
    assert np.array_equal(state.isFluid, host.isFluid)
    assert np.array_equal(state.masses, host.masses.astype(np.float32).astype(np.float64))
    assert np.abs(state.positions - host.positions).max() < 1e-12
    assert np.abs(state.velocities - host.velocities).max() < 1e-12
    assert np.abs(state.densities / host.densities - 1.0).max() < 1e-12
    lengths = host.lengths if host.lengths is not None else np.full(host.nParticles, case.smoothingLength)
    assert np.abs(state.lengths / lengths.astype(np.float32).astype(np.float64) - 1.0).max() < 1e-6
    batch.advance(10)
    assert int(batch.clocks()['steps'][0]) == steps + 10
    assert np.all(np.isfinite(batch.state(0).positions))


def testASplitThenMergedRunReportsTheDevicesMasses():
    # A split on the device followed by a merge of every daughter restores the run's count, but its
    # slots now hold merged particles in another order: state() must read masses, lengths and kinds
    # from the device, not the case's
    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.scenarios import solitaryOnSlope
    case = solitaryOnSlope(particlesPerDepth = 8, seed = 4)
    unsplit = float(case.particles.masses[case.particles.isFluid][0])
    batch = BatchEngine([case], 'float32', lanes = 2, listCapacity = 64)
    chosen = batch.split([(3.0, 5.0)], lengthRatio = 1.0)
    assert int(chosen[0]) > 0
    result = batch.merge([(3.0, 5.0)], unsplit, lengthRatio = 1.0)
    assert result[0].groups == int(chosen[0])
    assert batch.offsets[1] == case.count
    state = batch.state(0)
    device = batch.engine.download('mass', None, case.count)
    assert np.array_equal(state.masses, device)
    assert state.masses.sum() == pytest.approx(case.particles.masses.sum(), rel = 1e-6)
