'''

The single-run configuration's per-step derivative fields against the float64 restatement
(oracle/) on identical states: the 64-particles-per-depth solitary wave (85,707 particles, H/d and
slope drawn from seed 0), stepped 100 steps by the restatement, then the wall pressure, velocity
gradients G, correction matrices L, accelerations (limited viscosity) and continuity rate (L-corrected,
delta-SPH hydrostatic diffusion) taken on that state by the HIP engine's host steps in float32 pairs
and in the plain float64 kernel. Each field's relative L2 and L-infinity error (||d - o|| / ||o||,
max |d - o| / max |o|) is printed and held to 1e-4 in float32 and 1e-10 in float64.

'''

import numpy as np
import pytest

from oracle import oracle
from test_gpu_steps import physicsOf

pytestmark = pytest.mark.gpu

limits = {'float32': 1e-4, 'float64': 1e-10}


def errors(device: np.ndarray, expected: np.ndarray) -> tuple[float, float]:
    difference = np.asarray(device, dtype = float) - np.asarray(expected, dtype = float)
    return (float(np.linalg.norm(difference) / np.linalg.norm(expected)),
            float(np.abs(difference).max() / np.abs(expected).max()))


@pytest.mark.timeout(300)
@pytest.mark.parametrize('precision', ['float32', 'float64'])
def testTheC2DerivativeFieldsMatchTheOracle(precision):
    from pysurf_amd.pair_sums import PairSums
    from pysurf_amd.scenarios import singleRun
    case = singleRun(seed = 0, particlesPerDepth = 64)
    physics = physicsOf(case)
    p = case.particles
    run = oracle.OracleRun(physics, case.boundary.description(), p.positions, p.velocities, p.densities, p.masses,
                           p.isFluid)
    assert run.advance(np.inf, 100) == 100
    positions, velocities = run.positions.copy(), run.velocities.copy()
    densities, fluid, masses = run.dens[:run.count].copy(), p.isFluid, p.masses
    g = np.array([0.0, -9.80665])
    stiffness = case.referenceDensity * case.soundSpeed**2 / 7.0
    pressures = np.maximum(stiffness * ((densities / case.referenceDensity)**7 - 1.0), 0.0)
    walls, wallDensities = oracle.wallPressure(physics, positions, densities, pressures, masses, fluid)
    pressures = np.where(fluid, pressures, walls)
    gradients = oracle.velocityGradients(physics, positions, velocities, densities, masses, fluid)
    corrections = oracle.correctionMatrices(physics, positions, densities, masses, fluid)
    accelerations = oracle.accelerations(physics, positions, velocities, densities, pressures, masses, fluid,
                                         gradients = gradients)
    rates = oracle.continuityRate(physics, positions, velocities, densities, masses, fluid, corrections = corrections)

    sums = PairSums(2, case.smoothingLength, case.domainMin, case.domainMax, 0.0, 'float64', 'cuda:0', case.kernelType,
                    pairPrecision = precision)
    sums.update(positions, masses, fluid, None)
    sums.setDiffusion(case.diffusion, case.delta, case.soundSpeed, case.referenceDensity, 9.80665)
    sums.setViscosityLimiter(True)
    sums.setContinuityCorrection(True)
    fields = {
        'wall pressure': (sums.wallPressure(densities, pressures, g, case.referenceDensity, case.soundSpeed)[~fluid],
                          walls[~fluid]),
        'G': (sums.velocityGradients(velocities, densities), gradients),
        'L': (sums.correctionMatrices(densities), corrections),
        'acceleration': (sums.pairAccelerations(velocities, densities, pressures, case.soundSpeed, case.alpha,
                                                case.boundaryViscosityFactor, g), accelerations),
        'continuity': (sums.continuityRate(velocities, densities), rates),
    }
    for name, (device, expected) in fields.items():
        l2, linf = errors(device, expected)
        print(f'{precision} {name}: relative L2 {l2:.2e}, L-infinity {linf:.2e}', flush = True)
        assert l2 <= limits[precision] and linf <= limits[precision], name
