'''

The exact solitary wave of fully nonlinear potential flow and the δ-SPH start laid from it, after
pySurf's waveTheory.py:353-651 and waveMakers.py:328-442, 463-590: the wave by Clamond and Dutykh's
(2013) method, the flow under it from its conformal map, and the fluid particles, their velocities
and pressures, and the walls' pressures a run starts with, in 2D and over a 3D bed.

The method (Clamond & Dutykh, "Fast accurate computation of the fully nonlinear solitary surface
gravity waves", Computers & Fluids 84, 2013): a conformal map takes the water, in units of the
depth d and sqrt(g d), to the strip -1 <= beta <= 0 of alpha + i beta, the surface to beta = 0.
With C the operator of Fourier symbol k coth(k) (1 at k = 0), the surface elevation eta(alpha)
satisfies Babenko's equation

    F^2 eta - C^-1{eta} = C^-1{eta C{eta}} + eta^2 / 2,

solved by Petviashvili's iteration eta <- M^2 L^-1 N(eta), L = F^2 - C^-1 the linear part, N the
right-hand side and M = <eta, L eta> / <eta, N(eta)>, from the long-wave guess
(F^2 - 1) sech^2(kappa alpha / 2), F^2 = tan(kappa) / kappa. The map is z = zeta + X + i Y with Y
= eta on the surface and 0 on the bed; in the frame moving with the wave the flow is the strip's
uniform stream, so where the far water is at rest u - i v = F (1 - 1 / z'(zeta)).

'''

from __future__ import annotations

import functools
import math
from dataclasses import dataclass

import numpy as np

from pysurf_amd.particles import fillRegion, jitterPoints, latticePoints
from pysurf_amd.waves import gravity

froudeLimit = 1.29421          # the largest Froude number of a solitary wave (waveTheory.py:353)


@dataclass
class ExactSolitaryWave:

    '''

    The wave in units of d and sqrt(g d), moving toward +x over water at rest far from the crest
    (waveTheory.py:356-387): its Froude number c / sqrt(g d), amplitude H / d, and along the
    surface the positions from the crest x, the elevation eta and the potential (zero at the
    crest), at conformal coordinates alpha evenly spaced on a periodic interval, the crest at
    alpha = 0.

    '''

    froudeNumber: float
    amplitude: float
    x: np.ndarray
    eta: np.ndarray
    potential: np.ndarray
    alpha: np.ndarray


def wavenumbers(nodes: int, step: float) -> np.ndarray:
    return 2.0 * math.pi * np.fft.fftfreq(nodes, d = step)


def babenkoSolitaryWave(froudeNumber: float, nodes: int = 8192, tolerance: float = 1e-13,
                        maxIterations: int = 5000) -> ExactSolitaryWave:

    '''

    The exact solitary wave of Froude number `froudeNumber` in (1, froudeLimit) on `nodes` points
    (a power of two) of a periodic interval long enough for the starting guess to fall below
    1e-17, iterated until eta changes by less than `tolerance` (waveTheory.py:389-465). Raises
    ValueError outside the range and RuntimeError when the iteration does not settle.

    '''

    from scipy.optimize import brentq
    if not 1.0 < froudeNumber < froudeLimit:
        raise ValueError(f'Froude number {froudeNumber} is outside (1, {froudeLimit}).')
    F2 = froudeNumber * froudeNumber
    kappa = brentq(lambda k: math.tan(k) / k - F2, 1e-9, 0.5 * math.pi - 1e-9)
    # 4 (F^2 - 1) exp(-kappa L) below exp(-39.2), about 1e-17
    halfLength = (math.log(4.0 * (F2 - 1.0)) + 39.2) / kappa
    alpha = np.linspace(-halfLength, halfLength, nodes, endpoint = False)
    k = wavenumbers(nodes, 2.0 * halfLength / nodes)
    nonzero = k != 0.0
    coth = np.ones(nodes)
    coth[nonzero] = k[nonzero] / np.tanh(k[nonzero])          # the symbol of C
    tanhOver = 1.0 / coth                                      # the symbol of C^-1
    linear = F2 - tanhOver                                     # the symbol of L

    def filtered(symbol: np.ndarray, values: np.ndarray) -> np.ndarray:
        return np.fft.ifft(symbol * np.fft.fft(values)).real

    eta = (F2 - 1.0) / np.cosh(0.5 * kappa * alpha)**2
    change = float('inf')
    for _ in range(maxIterations):
        rhs = filtered(tanhOver, eta * filtered(coth, eta)) + 0.5 * eta * eta
        stabilizer = float(np.sum(eta * filtered(linear, eta)) / np.sum(eta * rhs))
        following = stabilizer * stabilizer * filtered(1.0 / linear, rhs)
        change = float(np.max(np.abs(following - eta)))
        eta = following
        if change < tolerance:
            break
    else:
        raise RuntimeError(f'The Babenko iteration did not settle for F = {froudeNumber}: last change {change:.2e}.')
    # X on the surface: its periodic part from eta's modes, -i eta_k coth(k), and eta's mean times alpha
    modes = np.fft.fft(eta)
    periodic = np.zeros(nodes, dtype = complex)
    periodic[nonzero] = -1j * modes[nonzero] / np.tanh(k[nonzero])
    shift = np.fft.ifft(periodic).real + float(np.mean(eta)) * alpha
    crest = nodes // 2
    shift -= shift[crest]
    return ExactSolitaryWave(froudeNumber = froudeNumber, amplitude = float(eta[crest]), x = alpha + shift, eta = eta,
                             potential = froudeNumber * shift, alpha = alpha)


@functools.lru_cache(maxsize = 32)
def exactSolitaryWave(heightRatio: float, nodes: int = 8192) -> ExactSolitaryWave:
    '''The exact wave of amplitude H / d (up to about 0.79), its Froude number found to 1e-13 (waveTheory.py:467-493).'''
    from scipy.optimize import brentq
    low = math.sqrt(1.0 + 0.8 * heightRatio)
    high = min(math.sqrt(1.0 + 1.05 * heightRatio), froudeLimit - 1e-9)
    froude = brentq(lambda F: babenkoSolitaryWave(F, nodes).amplitude - heightRatio, low, high, xtol = 1e-13)
    return babenkoSolitaryWave(froude, nodes)


def exactSolitaryWaveSpeed(height: float, depth: float, g: float = gravity) -> float:
    '''The exact wave's celerity [m/s] (waveTheory.py:495-499).'''
    return exactSolitaryWave(height / depth).froudeNumber * math.sqrt(g * depth)


def exactSolitarySurface(x: np.ndarray, height: float, depth: float, crestX: float = 0.0,
                         g: float = gravity) -> tuple[np.ndarray, np.ndarray]:
    '''Elevation [m] and surface potential [m^2/s] at increasing x, the potential zero at the first point (waveTheory.py:501-534).'''
    from scipy.interpolate import CubicSpline
    wave = exactSolitaryWave(height / depth)
    xi = np.clip((np.asarray(x, dtype = float) - crestX) / depth, wave.x[0], wave.x[-1])
    eta = CubicSpline(wave.x, wave.eta)(xi) * depth
    potential = CubicSpline(wave.x, wave.potential)(xi) * depth * math.sqrt(g * depth)
    return eta, potential - potential[0]


@functools.lru_cache(maxsize = 8)
def exactSolitaryInterior(heightRatio: float, levels: int = 129) -> tuple[np.ndarray, ...]:

    '''

    The water under the exact wave on `levels` conformal levels from the bed (beta = -1) to the
    surface (beta = 0), in units of d and sqrt(g d), the crest at x = 0 and still water at y = 0
    (waveTheory.py:536-594): x, y, u and v, each (levels, nodes), increasing in x along a level
    and in y from level to level.

    '''

    wave = exactSolitaryWave(heightRatio)
    alpha, F = wave.alpha, wave.froudeNumber
    nodes = len(alpha)
    k = wavenumbers(nodes, alpha[1] - alpha[0])
    nonzero = k != 0.0
    modes = np.fft.fft(wave.eta)
    mean = float(np.mean(wave.eta))
    kn = k[nonzero]

    def level(beta: float) -> tuple[np.ndarray, ...]:
        # Each mode of eta carried into the strip, and eta's mean as the linear part mean (zeta + i)
        spectrumX, spectrumY = np.zeros(nodes, dtype = complex), np.zeros(nodes, dtype = complex)
        spectrumX[nonzero] = -1j * modes[nonzero] * np.cosh(kn * (beta + 1.0)) / np.sinh(kn)
        spectrumY[nonzero] = modes[nonzero] * np.sinh(kn * (beta + 1.0)) / np.sinh(kn)
        X = np.fft.ifft(spectrumX).real + mean * alpha
        Y = np.fft.ifft(spectrumY).real + mean * (beta + 1.0)
        dX = np.fft.ifft(1j * k * spectrumX).real + mean
        dY = np.fft.ifft(1j * k * spectrumY).real
        return X, Y, dX, dY

    crestShift = level(0.0)[0][nodes // 2]
    x, y, u, v = (np.empty((levels, nodes)) for _ in range(4))
    for index, beta in enumerate(np.linspace(-1.0, 0.0, levels)):
        X, Y, dX, dY = level(beta)
        velocity = F * (1.0 - 1.0 / ((1.0 + dX) + 1j * dY))
        x[index], y[index] = alpha + X - crestShift, beta + Y
        u[index], v[index] = velocity.real, -velocity.imag
    return x, y, u, v


def exactSolitaryFlow(points: np.ndarray, height: float, depth: float, crestX: float = 0.0,
                      stillWaterLevel: float = 0.0, g: float = gravity,
                      chunk: int = 20000) -> tuple[np.ndarray, np.ndarray, np.ndarray]:

    '''

    Velocity [m/s], pressure over density [m^2/s^2] and the surface's elevation above still water
    at each point's x [m] under the exact wave over a flat bed, the far water at rest
    (waveTheory.py:596-651). The map's fields are interpolated along each level in x, then between
    the two levels around the point in y; a point above the surface takes the surface's flow, a
    point beyond the wave's interval is in still water. The wave keeps its form, so phi_t = -c u and
    unsteady Bernoulli gives p / rho = c u - |u|^2 / 2 - g (y - still water level). Points are
    taken `chunk` at a time.

    '''

    ratio = round(height / depth, 15)
    x, y, u, v = exactSolitaryInterior(ratio)
    froude = exactSolitaryWave(ratio).froudeNumber
    points = np.asarray(points, dtype = float).reshape(-1, 2)
    velocities, pressures, surfaces = np.zeros((len(points), 2)), np.zeros(len(points)), np.zeros(len(points))
    levels = x.shape[0]
    leftEdge, rightEdge = x[:, 0].max(), x[:, -1].min()
    for first in range(0, len(points), chunk):
        part = slice(first, first + chunk)
        xq = (points[part, 0] - crestX) / depth
        yq = (points[part, 1] - stillWaterLevel) / depth
        alongY = np.vstack([np.interp(xq, x[l], y[l]) for l in range(levels)])
        alongU = np.vstack([np.interp(xq, x[l], u[l]) for l in range(levels)])
        alongV = np.vstack([np.interp(xq, x[l], v[l]) for l in range(levels)])
        below = np.clip(np.sum(alongY <= yq, axis = 0) - 1, 0, levels - 2)
        columns = np.arange(len(xq))
        lowY, highY = alongY[below, columns], alongY[below + 1, columns]
        share = np.clip((yq - lowY) / (highY - lowY), 0.0, 1.0)
        uq = (1.0 - share) * alongU[below, columns] + share * alongU[below + 1, columns]
        vq = (1.0 - share) * alongV[below, columns] + share * alongV[below + 1, columns]
        surface = alongY[-1].copy()
        still = (xq <= leftEdge) | (xq >= rightEdge)
        uq[still], vq[still], surface[still] = 0.0, 0.0, 0.0
        velocities[part] = np.column_stack([uq, vq]) * math.sqrt(g * depth)
        pressures[part] = (froude * uq - 0.5 * (uq * uq + vq * vq) - yq) * g * depth
        surfaces[part] = surface * depth
    return velocities, pressures, surfaces


def exactSurfaceElevation(x: np.ndarray, height: float, depth: float, crestX: float) -> np.ndarray:
    '''The surface exactSolitaryFlow reads, above still water [m], without the flow (waveMakers.py:328-343).'''
    levelX, levelY, _, _ = exactSolitaryInterior(round(height / depth, 15))
    xq = (np.asarray(x, dtype = float) - crestX) / depth
    surface = np.interp(xq, levelX[-1], levelY[-1])
    still = (xq <= levelX[:, 0].max()) | (xq >= levelX[:, -1].min())
    return np.where(still, 0.0, surface) * depth


def bedFittedPoints(lower, upper, spacing: float, floor) -> np.ndarray:

    '''

    Columns standing on a floor (sphParticles.py:314-355): at the lattice's horizontal positions in
    [lower, upper], each column from half a spacing above floor(horizontal...) (or the box's
    bottom, where the floor lies below it) up to the box's top, a spacing apart, counted as
    np.arange counts. Over a flat floor at the bottom they are the lattice's points.

    '''

    lower, upper = np.asarray(lower, dtype = float), np.asarray(upper, dtype = float)
    axes = [np.arange(low + spacing / 2.0, high, spacing) for low, high in zip(lower[:-1], upper[:-1])]
    horizontal = [grid.ravel() for grid in np.meshgrid(*axes, indexing = 'ij')]
    base = np.maximum(np.asarray(floor(*horizontal), dtype = float), lower[-1]) + spacing / 2.0
    counts = np.maximum(np.ceil((upper[-1] - base) / spacing), 0.0).astype(int)
    owner = np.repeat(np.arange(base.size), counts)
    row = np.arange(owner.size) - np.repeat(np.cumsum(counts) - counts, counts)
    return np.column_stack([coordinate[owner] for coordinate in horizontal] + [base[owner] + row * spacing])


def exactSolitaryWaveFluid(bathymetry, height: float, crestX: float, spacing: float, bedFitted: bool = False,
                           jitter: float = 0.0, seed: int = 1) -> tuple[np.ndarray, np.ndarray, np.ndarray]:

    '''

    Fluid particles under the exact wave over a bed, with their velocities and pressures over
    density (waveMakers.py:345-420): points from half a spacing above the bed up to the wave's
    surface, on the lattice or in bed-fitted columns, moved by up to `jitter` spacings with `seed`
    when asked; the flow is that of a flat bed of the depth at the crest, evaluated at each
    particle's final place, the pressure floored at zero.

    '''

    depth = float(bathymetry.depth(crestX))
    level = bathymetry.stillWaterLevel

    def inside(points: np.ndarray) -> np.ndarray:
        bed = bathymetry.elevation(points[:, 0])
        surface = exactSurfaceElevation(points[:, 0], height, depth, crestX)
        # The tolerance keeps the lattice row exactly half a spacing above a flat bed
        return (points[:, 1] >= bed + (0.5 - 1e-9) * spacing) & (points[:, 1] < level + surface)

    lower = [bathymetry.startX, float(np.min(bathymetry.z))]
    upper = [bathymetry.endX, level + height + spacing]
    if bedFitted:
        candidates = bedFittedPoints(lower, upper, spacing, bathymetry.elevation)
        positions = candidates[inside(candidates)]
    else:
        positions = fillRegion(inside, lower, upper, spacing)
    if jitter > 0.0:
        surface = lambda x: level + exactSurfaceElevation(x, height, depth, crestX)
        positions = jitterPoints(positions, spacing, jitter, seed, [bathymetry.startX], [bathymetry.endX],
                                 bathymetry.elevation, surface)
    velocities, pressures, _ = exactSolitaryFlow(positions, height, depth, crestX, level)
    return positions, velocities, np.maximum(pressures, 0.0)


def exactSolitaryBoundaryPressure(bathymetry, positions: np.ndarray, height: float, crestX: float, spacing: float,
                                  g: float = gravity) -> np.ndarray:

    '''

    Each wall particle's starting pressure over density under the exact wave (waveMakers.py:422-442):
    the water's at the nearest point half a spacing inside it (above the bed, in from the end
    walls) plus the hydrostatic increment down to the particle; zero where that point is dry.

    '''

    depth = float(bathymetry.depth(crestX))
    level = bathymetry.stillWaterLevel
    positions = np.asarray(positions, dtype = float)
    inner = np.column_stack([np.clip(positions[:, 0], bathymetry.startX + 0.5 * spacing, bathymetry.endX - 0.5 * spacing),
                             positions[:, 1]])
    inner[:, 1] = np.maximum(inner[:, 1], bathymetry.elevation(inner[:, 0]) + 0.5 * spacing)
    _, pressure, surface = exactSolitaryFlow(inner, height, depth, crestX, level, g)
    wet = inner[:, 1] < level + surface
    return np.where(wet, np.maximum(pressure, 0.0) + g * (inner[:, 1] - positions[:, 1]), 0.0)


def exactSolitaryWaveFluid3D(field, height: float, depth: float, crestX: float, spacing: float,
                             floor: float | None = None, bedFitted: bool = False, jitter: float = 0.0,
                             seed: int = 1) -> tuple[np.ndarray, np.ndarray, np.ndarray]:

    '''

    Fluid particles under the exact wave, its crest straight along y, over a 3D bed (a HeightField),
    with their velocities and pressures over density (waveMakers.py:463-568): the flat-bed wave of
    `depth` in every (x, z) plane, no spanwise velocity. Points inside the field's bounds from half
    a spacing above the bed up to the surface, on the lattice counted from `floor` (the bed's lowest
    point by default) or in bed-fitted columns; with `jitter`, moved by up to that many spacings
    with `seed`, and the flow taken at each point's place.

    '''

    s, level = spacing, field.stillWaterLevel
    floor = field.lowestElevation() if floor is None else floor
    startX, endX, startY, endY = field.bounds
    top = level + height + s
    surfaceAt = lambda x: exactSurfaceElevation(x, height, depth, crestX)
    if jitter > 0.0:
        if bedFitted:
            points = bedFittedPoints([startX, startY, floor], [endX, endY, top], s, field.elevation)
        else:
            points = latticePoints([startX, startY, floor], [endX, endY, top], s)
            points = points[points[:, 2] >= field.elevation(points[:, 0], points[:, 1]) + (0.5 - 1e-9) * s]
        points = points[points[:, 2] < level + surfaceAt(points[:, 0])]
        points = jitterPoints(points, s, jitter, seed, [startX, startY], [endX, endY],
                              lambda x, y: np.maximum(field.elevation(x, y), floor), lambda x, y: level + surfaceAt(x))
        velocities, pressures, _ = exactSolitaryFlow(points[:, [0, 2]], height, depth, crestX, level)
        flow = np.column_stack([velocities[:, 0], np.zeros(len(points)), velocities[:, 1]])
        return points, flow, np.maximum(pressures, 0.0)
    if bedFitted:
        # The flow once per distinct (x, z): rows across a bed flat in y share theirs
        points = bedFittedPoints([startX, startY, floor], [endX, endY, top], s, field.elevation)
        planar, inverse = np.unique(points[:, [0, 2]], axis = 0, return_inverse = True)
        velocities, pressures, surface = exactSolitaryFlow(planar, height, depth, crestX, level)
        inverse = inverse.reshape(-1)
        wet = points[:, 2] < level + surface[inverse]
        index = inverse[wet]
        flow = np.column_stack([velocities[index, 0], np.zeros(len(index)), velocities[index, 1]])
        return points[wet], flow, np.maximum(pressures[index], 0.0)
    # The flow on one (x, z) plane of the lattice, laid at every row across
    plane = latticePoints([startX, floor], [endX, top], s)
    velocities, pressures, surface = exactSolitaryFlow(plane, height, depth, crestX, level)
    wet = plane[:, 1] < level + surface
    plane, velocities, pressures = plane[wet], velocities[wet], np.maximum(pressures[wet], 0.0)
    rows = np.arange(startY + 0.5 * s, endY, s)
    x, z = np.tile(plane[:, 0], rows.size), np.tile(plane[:, 1], rows.size)
    y = np.repeat(rows, len(plane))
    keep = z >= field.elevation(x, y) + (0.5 - 1e-9) * s
    index = np.tile(np.arange(len(plane)), rows.size)[keep]
    flow = np.column_stack([velocities[index, 0], np.zeros(len(index)), velocities[index, 1]])
    return np.column_stack([x[keep], y[keep], z[keep]]), flow, pressures[index]


def exactSolitaryBoundaryPressure3D(field, positions: np.ndarray, height: float, depth: float, crestX: float,
                                    spacing: float, g: float = gravity) -> np.ndarray:
    '''A 3D tank's wall particles' starting pressure over density under the exact wave, by the 2D rule, in from the side walls too (waveMakers.py:570-590).'''
    startX, endX, startY, endY = field.bounds
    level = field.stillWaterLevel
    positions = np.asarray(positions, dtype = float)
    innerX = np.clip(positions[:, 0], startX + 0.5 * spacing, endX - 0.5 * spacing)
    innerY = np.clip(positions[:, 1], startY + 0.5 * spacing, endY - 0.5 * spacing)
    innerZ = np.maximum(positions[:, 2], field.elevation(innerX, innerY) + 0.5 * spacing)
    _, pressure, surface = exactSolitaryFlow(np.column_stack([innerX, innerZ]), height, depth, crestX, level, g)
    wet = innerZ < level + surface
    return np.where(wet, np.maximum(pressure, 0.0) + g * (innerZ - positions[:, 2]), 0.0)


__all__ = ['ExactSolitaryWave', 'babenkoSolitaryWave', 'exactSolitaryWave', 'exactSolitaryWaveSpeed',
           'exactSolitarySurface', 'exactSolitaryInterior', 'exactSolitaryFlow', 'exactSurfaceElevation',
           'bedFittedPoints', 'exactSolitaryWaveFluid', 'exactSolitaryBoundaryPressure', 'exactSolitaryWaveFluid3D',
           'exactSolitaryBoundaryPressure3D']
