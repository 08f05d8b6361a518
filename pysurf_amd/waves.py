'''

The first-order solitary wave the configurations start from: celerity, decay, profile and the
long-wave velocity field, after pySurf's waveTheory.py:192-259 and waveMakers.py:242-326.

'''

from __future__ import annotations

import math

import numpy as np

from pysurf_amd.particles import fillRegion

gravity = 9.80665             # constants.py:31
freshwaterDensity = 998.2     # constants.py:34


def solitaryWaveSpeed(height: float, depth: float, g: float = gravity) -> float:
    '''c = sqrt(g (d + H)) (waveTheory.py:192-196).'''
    return math.sqrt(g * (depth + height))


def solitaryWaveDecay(height: float, depth: float) -> float:
    '''kappa = sqrt(3 H / (4 d^3)) (waveTheory.py:198-202).'''
    return math.sqrt(3.0 * height / (4.0 * depth**3))


def solitaryWaveProfile(x, height: float, depth: float, crestX: float = 0.0) -> np.ndarray:
    '''eta = H sech^2(kappa (x - x0)) (waveTheory.py:204-227).'''
    argument = solitaryWaveDecay(height, depth) * (np.asarray(x, dtype = float) - crestX)
    return height / np.cosh(argument)**2


def solitaryWaveHalfLength(height: float, depth: float, fraction: float = 0.01) -> float:
    '''Distance from the crest where eta falls to `fraction` of H (waveTheory.py:245-259).'''
    return math.acosh(1.0 / math.sqrt(fraction)) / solitaryWaveDecay(height, depth)


def solitaryWaveVelocityField(x, y, height: float, depth: float, crestX: float, bedElevation) -> np.ndarray:
    '''u = c eta / (d + eta), w = -(y - z_b) c d eta' / (d + eta)^2 (waveMakers.py:242-280).'''
    celerity = solitaryWaveSpeed(height, depth)
    kappa = solitaryWaveDecay(height, depth)
    eta = solitaryWaveProfile(x, height, depth, crestX)
    etaSlope = -2.0 * kappa * eta * np.tanh(kappa * (x - crestX))
    u = celerity * eta / (depth + eta)
    w = -(y - bedElevation) * celerity * depth * etaSlope / (depth + eta)**2
    return np.column_stack([u, w])


def solitaryWaveFluid(bathymetry, height: float, crestX: float, spacing: float) -> tuple[np.ndarray, np.ndarray]:
    '''Lattice points from half a spacing above the bed to SWL + eta, and their velocities (waveMakers.py:282-326).'''
    depth = float(bathymetry.depth(crestX))
    swl = bathymetry.stillWaterLevel

    def inside(points):
        bed = bathymetry.elevation(points[:, 0])
        surface = swl + solitaryWaveProfile(points[:, 0], height, depth, crestX)
        return (points[:, 1] >= bed + (0.5 - 1e-9) * spacing) & (points[:, 1] < surface)

    lower = [bathymetry.startX, float(np.min(bathymetry.z))]
    upper = [bathymetry.endX, swl + height + spacing]
    positions = fillRegion(inside, lower, upper, spacing)
    velocities = solitaryWaveVelocityField(positions[:, 0], positions[:, 1], height, depth, crestX,
                                           bathymetry.elevation(positions[:, 0]))
    return positions, velocities


def solitaryWaveFluid3D(field, height: float, crestX: float, spacing: float) -> tuple[np.ndarray, np.ndarray]:

    '''

    The 3D start over a height field: lattice points from half a spacing above the bed to
    SWL + eta(x), the crest line along y, with the long-wave velocity of the 2D start in each x-z
    plane (waveMakers.py:282-326 extended across y as BreakingWave3D lays its wave).

    '''

    depth = field.stillWaterLevel - float(field.elevation(field.bounds[0], 0.5 * (field.bounds[2] + field.bounds[3])))
    swl = field.stillWaterLevel
    startX, endX, startY, endY = field.bounds

    def inside(points):
        bed = field.elevation(points[:, 0], points[:, 1])
        surface = swl + solitaryWaveProfile(points[:, 0], height, depth, crestX)
        return (points[:, 2] >= bed + (0.5 - 1e-9) * spacing) & (points[:, 2] < surface)

    lower = [startX, startY, field.lowestElevation()]
    upper = [endX, endY, swl + height + spacing]
    positions = fillRegion(inside, lower, upper, spacing)
    planar = solitaryWaveVelocityField(positions[:, 0], positions[:, 2], height, depth, crestX,
                                       field.elevation(positions[:, 0], positions[:, 1]))
    velocities = np.column_stack([planar[:, 0], np.zeros(len(positions)), planar[:, 1]])
    return positions, velocities
