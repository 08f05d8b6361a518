'''

The solver shell's handling of the options the delta-SPH scheme leaves off (sphSolver.py:320-380):
the device step carries the Shepard filter, the 'antuono' diffusion and the closure, but not XSPH
or shifting, so WcsphSolver refuses those two before touching a device, naming the PairSums host
step that gives each. CPU only.

'''

import numpy as np
import pytest

from pysurf_amd.solver import SimulationConfig, WcsphSolver


def deltaSph(**changes) -> SimulationConfig:
    '''The delta-SPH scheme's settings (BreakingWave.py:722-725) with `changes` applied.'''
    settings = dict(domainMin = np.zeros(2), domainMax = np.array([1.0, 0.5]), particleSpacing = 0.02,
                    smoothingLengthRatio = 3.0, kernelType = 'wendlandC2', xsphEpsilon = 0.0,
                    densityFilterInterval = 0, densityDiffusion = 'hydrostatic', wallPressure = 'extrapolated',
                    continuityCorrection = True, viscosityLimiter = True)
    settings.update(changes)
    return SimulationConfig(**settings)


@pytest.mark.parametrize('changes, method', [
    (dict(xsphEpsilon = 0.5), 'xsphVelocity'),
    (dict(particleShifting = True), 'shiftingDisplacement'),
])
def testTheSolverRefusesEachHostOnlyOption(changes, method):
    with pytest.raises(ValueError, match = method):
        WcsphSolver(deltaSph(**changes))


def testTheSolverStillRefusesUnknownOptions():
    with pytest.raises(ValueError, match = 'closure'):
        WcsphSolver(deltaSph(turbulenceClosure = 'smagorinsky'))
    with pytest.raises(ValueError, match = 'diffusion'):
        WcsphSolver(deltaSph(densityDiffusion = 'laplacian'))
    with pytest.raises(ValueError, match = 'filter interval'):
        WcsphSolver(deltaSph(densityFilterInterval = -1))
