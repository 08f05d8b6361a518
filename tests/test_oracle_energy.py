'''

The float64 restatement's energy budget and resident options (oracle/sph_oracle.cpp), checked
without the device:

  - the elastic energy e(rho) against a quadrature of its definition, the integral of p / s^2
    from rho0 (sphEnergy.py:116-160), with and without tension;
  - the ledger's 14 terms against the change of the fluid's K + P + E_C over steps of the
    delta-SPH scheme, with the Shepard filter, the 'antuono' diffusion and the delta-LES-SPH
    closure in both of the reference's forms (tests/testSphEnergy.py:201-259: the budget closes);
  - the viscous term never adds energy, the correction's term is zero without the correction and
    the filter's is zero without the filter (tests/testSphEnergy.py);
  - a filtered step is the Shepard density (sphGpu.py:2251-2254) followed by an unfiltered step;
  - still water stays at rest under each option (tests/testSphGpu.py:557-575, 620-637: the
    fastest particle under 1% of sqrt(g d) and no particle moved a quarter spacing after 0.2 s).

A slip in any term (a sign, a factor 2, a dropped part) breaks the closure, which sums every term
against an energy computed from the state alone.

'''

import math

import numpy as np
import pytest
from scipy import integrate

from oracle import oracle

closureForms = {'published': dict(diffusion = 'antuono', alpha = 0.0, limiter = False),
                'hybrid': dict(diffusion = 'antuono', alpha = 0.01, limiter = True)}
closureFluid = dict(soundSpeed = None, referenceDensity = 1000.0, kinematicViscosity = 1e-6)


def caseFor(ppd: int = 6, waveHeight: float = 0.45, seed: int = 0):
    from pysurf_amd.scenarios import solitaryOnSlope
    return solitaryOnSlope(particlesPerDepth = ppd, waveHeight = waveHeight, seed = seed)


def physicsOf(case, **changes) -> oracle.Physics:
    options = dict(dims = case.dimensions, h = case.smoothingLength, domainMin = case.domainMin,
                   domainMax = case.domainMax, kernelType = case.kernelType, rho0 = case.referenceDensity,
                   c0 = case.soundSpeed, alpha = case.alpha, boundaryFactor = case.boundaryViscosityFactor,
                   gravity = case.gravity, diffusion = case.diffusion, delta = case.delta,
                   extrapolatedWalls = case.extrapolatedWalls, negativePressure = case.negativePressure,
                   correction = case.correction, limiter = case.limiter, densityFloor = case.densityFloor,
                   cfl = case.cfl, maxTimeStep = case.maxTimeStep)
    options.update(changes)
    return oracle.Physics(**options)


def runOf(case, physics, **options) -> oracle.OracleRun:
    p = case.particles
    return oracle.OracleRun(physics, case.boundary.description(), p.positions, p.velocities, p.densities, p.masses,
                            p.isFluid, lengths = p.lengths, **options)


def closureOf(case) -> dict:
    return dict(closureFluid, soundSpeed = case.soundSpeed)


@pytest.mark.parametrize('tension', [False, True])
def testTheElasticEnergyIsTheIntegralOfPressureOverDensitySquared(tension):
    case = caseFor()
    physics = physicsOf(case, negativePressure = tension)
    rho0, c0 = physics.rho0, physics.c0
    stiffness = rho0 * c0 * c0 / 7.0

    def pressure(rho):
        value = stiffness * ((rho / rho0) ** 7 - 1.0)
        return value if tension else max(value, 0.0)

    densities = rho0 * np.array([0.9, 0.97, 0.999, 1.0, 1.0001, 1.003, 1.02, 1.08])
    energies = oracle.elasticEnergy(physics, densities)
    for rho, e in zip(densities, energies):
        expected, _ = integrate.quad(lambda s: pressure(s) / (s * s), rho0, rho, epsabs = 0.0, epsrel = 1e-13)
        assert e == pytest.approx(expected, rel = 1e-10, abs = 1e-14 * c0 * c0), rho
    # near rho0 it is linear acoustics' c0^2 x^2 / 2
    x = 1e-4
    assert oracle.elasticEnergy(physics, [rho0 * (1.0 + x)])[0] == pytest.approx(c0 * c0 * x * x / 2.0, rel = 1e-3)


def throughput(terms: dict) -> float:
    return sum(abs(v) for v in terms.values())


variants = {
    'delta-SPH': (dict(), dict()),
    'filtered': (dict(), dict(filterInterval = 4)),
    'antuono': (dict(diffusion = 'antuono', correction = False), dict()),
    'closure-published': (closureForms['published'], dict(closure = True)),
    'closure-hybrid': (closureForms['hybrid'], dict(closure = True)),
}


@pytest.mark.parametrize('variant', list(variants))
def testTheLedgerClosesOnTheFluidEnergy(variant):
    changes, options = variants[variant]
    case = caseFor()
    physics = physicsOf(case, **changes)
    if options.get('closure'):
        options = dict(options, closure = closureOf(case))
    run = runOf(case, physics, ledger = True, **options)
    start = run.fluidEnergy()
    assert run.advance(np.inf, 30) == 30
    budget = run.budget()
    change = run.fluidEnergy() - start
    assert sum(budget.values()) == pytest.approx(change, abs = 1e-9 * throughput(budget))
    assert budget['viscosityFluid'] <= 0.0
    assert budget['shifting'] == 0.0
    if not physics.correction:
        assert budget['correction'] == 0.0
    if not options.get('filterInterval'):
        assert budget['filter'] == 0.0
    else:
        assert budget['filter'] != 0.0
    # the step does move energy between the terms
    assert budget['pressureFluid'] != 0.0 and budget['gravityWork'] != 0.0 and budget['drift'] != 0.0


def testTheLedgerLeavesTheStepAlone():
    case = caseFor(seed = 2)
    plain = runOf(case, physicsOf(case))
    tracked = runOf(case, physicsOf(case), ledger = True)
    plain.advance(np.inf, 20)
    tracked.advance(np.inf, 20)
    for name in ('pos', 'vel', 'dens'):
        assert np.array_equal(getattr(plain, name), getattr(tracked, name)), name


def testAFilteredStepIsTheShepardDensityThenTheStep():
    case = caseFor(seed = 3)
    physics = physicsOf(case)
    filtered = runOf(case, physics, filterInterval = 5)
    assert filtered.advance(np.inf, 6) == 6
    composed = runOf(case, physics)
    composed.advance(np.inf, 5)
    fluid = composed.fluid != 0
    shepard = oracle.shepardDensity(physics, composed.positions, composed.dens, composed.mass, fluid,
                                    lengths = composed.lens)
    composed.dens[fluid] = shepard[fluid]
    composed.advance(np.inf, 1)
    for name in ('pos', 'vel', 'dens'):
        assert np.array_equal(getattr(filtered, name), getattr(composed, name)), name
    # and the filter's first use is at step 5, not at the start
    early = runOf(case, physics, filterInterval = 5)
    early.advance(np.inf, 5)
    plain = runOf(case, physics)
    plain.advance(np.inf, 5)
    assert np.array_equal(early.dens, plain.dens)


# The reference's still tank (tests/testSphGpu.py:440-457, 554-575, 602-637): alpha 0.1 without the
# limiter under the delta-SPH continuity, and the closure's two forms
tankForms = {
    'hydrostatic': (dict(alpha = 0.1, limiter = False, diffusion = 'hydrostatic'), dict()),
    'antuono': (dict(alpha = 0.1, limiter = False, diffusion = 'antuono'), dict()),
    'filtered': (dict(alpha = 0.1, limiter = False, diffusion = 'hydrostatic'), dict(filterInterval = 30)),
    'closure-published': (dict(closureForms['published']), dict(closure = True)),
    'closure-hybrid': (dict(closureForms['hybrid']), dict(closure = True)),
}


def tankCase(form: str):
    from pysurf_amd.scenarios import stillTank
    physics, options = tankForms[form]
    case = stillTank(maxTimeStep = 1e-3, **physics)
    if options.get('closure'):
        options = dict(options, closure = closureOf(case))
    return case, options


@pytest.mark.parametrize('form', list(tankForms))
def testStillWaterStaysAtRest(form):
    case, options = tankCase(form)
    run = runOf(case, physicsOf(case), **options)
    fluid = run.fluid != 0
    start = run.pos[fluid].copy()
    run.advance(0.2, 100000)
    assert run.time.value >= 0.2
    depth = case.parameters['depth']
    speed = np.sqrt(np.sum(run.vel[fluid] ** 2, axis = 1)).max()
    assert speed < 0.01 * math.sqrt(9.81 * depth)
    assert np.abs(run.pos[fluid] - start).max() < 0.25 * case.spacing
