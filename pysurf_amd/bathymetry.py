'''

Bed profiles the boundary and the start are laid on: pySurf's `Bathymetry` polyline
(src/pysurf/bathymetry.py:66-308) and its `HeightField` (bathymetry.py:370-850), the parts the
delta-SPH runs use.

'''

from __future__ import annotations

import math

import numpy as np


class Bathymetry:

    '''A 2D bed profile, piecewise linear through its vertices, still water at `stillWaterLevel` (bathymetry.py:66-96).'''

    def __init__(self, x, z, stillWaterLevel: float) -> None:
        self.x = np.asarray(x, dtype = float)
        self.z = np.asarray(z, dtype = float)
        self.stillWaterLevel = float(stillWaterLevel)
        if self.x.ndim != 1 or self.x.size < 2 or self.z.shape != self.x.shape:
            raise ValueError('A bed profile needs at least two (x, z) vertices of matching length.')
        if not (np.all(np.isfinite(self.x)) and np.all(np.isfinite(self.z))):
            raise ValueError('Bed vertices must be finite.')
        if np.any(np.diff(self.x) <= 0.0):
            raise ValueError('Bed vertex x must increase strictly, so the bed is a function of x.')

    @classmethod
    def planeBeach(cls, depth: float, toeX: float, slope: float, endX: float, startX: float = 0.0) -> Bathymetry:
        '''Flat at z = 0 to the toe, then a constant slope to endX; still water at z = depth (bathymetry.py:115-144).'''
        if not (startX < toeX < endX):
            raise ValueError('A plane beach needs startX < toeX < endX.')
        if slope <= 0.0 or depth <= 0.0:
            raise ValueError('Depth and slope must be positive.')
        return cls([startX, toeX, endX], [0.0, 0.0, (endX - toeX) * slope], stillWaterLevel = depth)

    @property
    def startX(self) -> float:
        return float(self.x[0])

    @property
    def endX(self) -> float:
        return float(self.x[-1])

    def elevation(self, x) -> np.ndarray:
        '''z_b(x), constant beyond the ends (bathymetry.py:221-225).'''
        return np.interp(x, self.x, self.z)

    def depth(self, x) -> np.ndarray:
        '''max(0, SWL - z_b(x)) (bathymetry.py:227-231).'''
        return np.maximum(0.0, self.stillWaterLevel - self.elevation(x))

    def slope(self, x) -> np.ndarray:
        '''dz_b/dx of the segment holding x (bathymetry.py:233-240).'''
        segmentSlopes = np.diff(self.z) / np.diff(self.x)
        index = np.clip(np.searchsorted(self.x, x, side = 'right') - 1, 0, segmentSlopes.size - 1)
        return segmentSlopes[index]

    def normal(self, x) -> np.ndarray:
        '''Unit normal up out of the bed, shape (..., 2) (bathymetry.py:242-250).'''
        slope = np.asarray(self.slope(x), dtype = float)
        length = np.sqrt(1.0 + slope**2)
        return np.stack([-slope / length, 1.0 / length], axis = -1)

    def sample(self, spacing: float) -> tuple[np.ndarray, np.ndarray]:
        '''Points along the bed at equal arc length, both ends included (bathymetry.py:276-297).'''
        segmentLengths = np.hypot(np.diff(self.x), np.diff(self.z))
        arcLength = np.concatenate([[0.0], np.cumsum(segmentLengths)])
        count = max(2, int(round(arcLength[-1] / spacing)) + 1)
        targets = np.linspace(0.0, arcLength[-1], count)
        return np.interp(targets, arcLength, self.x), np.interp(targets, arcLength, self.z)


def convolutionWeights(u) -> np.ndarray:
    '''Cubic convolution weights (Keys, a = -1/2) of the 4 grid points around fraction u, shape (4, ...) (bathymetry.py:310-324).'''
    u = np.asarray(u, dtype = float)
    u2, u3 = u * u, u * u * u
    return np.stack([-0.5 * u3 + u2 - 0.5 * u, 1.5 * u3 - 2.5 * u2 + 1.0, -1.5 * u3 + 2.0 * u2 + 0.5 * u,
                     0.5 * u3 - 0.5 * u2])


def convolutionSlopeWeights(u) -> np.ndarray:
    '''The u-derivatives of convolutionWeights, shape (4, ...) (bathymetry.py:326-333).'''
    u = np.asarray(u, dtype = float)
    u2 = u * u
    return np.stack([-1.5 * u2 + 2.0 * u - 0.5, 4.5 * u2 - 5.0 * u, -4.5 * u2 + 4.0 * u + 0.5, 1.5 * u2 - u])


class HeightField:

    '''

    A 3D bed z_b(x, y) on a regular grid, interpolated by cubic convolution from the 4 x 4 grid
    points around a query, so the surface and its gradient are continuous and quadratics are
    reproduced exactly; past the grid's edge a query takes the edge's value and a zero gradient
    across it (bathymetry.py:370-480). x runs along the waves, y across them, z up; `bounds`
    (startX, endX, startY, endY) is the region the bed covers, the tank's walls.

    '''

    def __init__(self, x, y, z, stillWaterLevel: float, bounds: tuple | None = None) -> None:
        self.x = np.asarray(x, dtype = float)
        self.y = np.asarray(y, dtype = float)
        self.z = np.asarray(z, dtype = float)
        self.stillWaterLevel = float(stillWaterLevel)
        if self.x.ndim != 1 or self.y.ndim != 1 or self.x.size < 4 or self.y.size < 4:
            raise ValueError('A height field needs at least four grid points along x and y.')
        if self.z.shape != (self.y.size, self.x.size):
            raise ValueError('The elevations must be shaped (len(y), len(x)).')
        if not (np.all(np.isfinite(self.x)) and np.all(np.isfinite(self.y)) and np.all(np.isfinite(self.z))):
            raise ValueError('Height field values must be finite.')
        for axis in (self.x, self.y):
            steps = np.diff(axis)
            if np.any(steps <= 0.0) or not np.allclose(steps, steps[0], rtol = 1e-9, atol = 0.0):
                raise ValueError('The grid must increase in equal steps along each axis.')
        if bounds is None:
            bounds = (self.x[0], self.x[-1], self.y[0], self.y[-1])
        self.bounds = tuple(float(value) for value in bounds)
        startX, endX, startY, endY = self.bounds
        if not (self.x[0] - 1e-9 <= startX < endX <= self.x[-1] + 1e-9
                and self.y[0] - 1e-9 <= startY < endY <= self.y[-1] + 1e-9):
            raise ValueError('The bounds must lie inside the grid.')

    @classmethod
    def fromFunction(cls, function, bounds: tuple, gridSpacing: float, stillWaterLevel: float,
                     padding: int = 2) -> HeightField:
        '''An analytic bed sampled over its bounds and `padding` cells beyond them, each axis evenly divided (bathymetry.py:481-516).'''
        startX, endX, startY, endY = (float(value) for value in bounds)
        if not (endX > startX and endY > startY and gridSpacing > 0.0):
            raise ValueError('A height field needs positive extents and grid spacing.')
        axes = []
        for low, high in ((startX, endX), (startY, endY)):
            cells = max(1, int(math.ceil((high - low) / gridSpacing - 1e-9)))
            step = (high - low) / cells
            axes.append(low + step * np.arange(-padding, cells + padding + 1))
        gridX, gridY = np.meshgrid(*axes, indexing = 'xy')
        return cls(axes[0], axes[1], np.asarray(function(gridX, gridY), dtype = float), stillWaterLevel,
                   bounds = (startX, endX, startY, endY))

    @classmethod
    def sechRidge(cls, k: float, depth: float = 1.0, toeX: float = 5.225, slope: float = 1.0 / 15.0,
                  endX: float = 19.0, halfWidth: float = 4.0, startX: float = 0.0,
                  gridSpacing: float | None = None) -> HeightField:
        '''The sloping ridge z_b / d = slope (x / d - toeX) sech^2(k y / d) shoreward of the toe (bathymetry.py:518-553).'''

        def elevation(x, y):
            return depth * slope * np.maximum(x / depth - toeX, 0.0) / np.cosh(k * y / depth)**2

        return cls.fromFunction(elevation, (startX * depth, endX * depth, -halfWidth * depth, halfWidth * depth),
                                gridSpacing if gridSpacing is not None else depth / 64.0, stillWaterLevel = depth)

    @classmethod
    def gaussianRidge(cls, depth: float = 1.0, height: float = 0.5, width: float = 2.0, slope: float = 1.0 / 15.0,
                      toeX: float = 6.0, length: float = 20.0, halfWidth: float = 5.0, direction: float = 25.0,
                      gridSpacing: float | None = None) -> HeightField:

        '''

        A ridge of Gaussian cross-section whose crest rises shoreward from its toe at `slope` until
        it stands `height` above the flat bed, its axis turned `direction` degrees from x about
        (toeX, 0), in a tank [0, length] x [-halfWidth, halfWidth], all in units of `depth`:

            z_b / d = min(slope max(xi - toeX, 0), height) exp(-eta^2 / (2 width^2)),
            xi = toeX + (x - toeX) cos(t) + y sin(t),   eta = -(x - toeX) sin(t) + y cos(t).

        '''

        turn = math.radians(direction)

        def elevation(x, y):
            along = toeX + (x / depth - toeX) * math.cos(turn) + (y / depth) * math.sin(turn)
            across = -(x / depth - toeX) * math.sin(turn) + (y / depth) * math.cos(turn)
            crest = np.minimum(slope * np.maximum(along - toeX, 0.0), height)
            return depth * crest * np.exp(-across**2 / (2.0 * width**2))

        return cls.fromFunction(elevation, (0.0, length * depth, -halfWidth * depth, halfWidth * depth),
                                gridSpacing if gridSpacing is not None else depth / 64.0, stillWaterLevel = depth)

    @property
    def spacingX(self) -> float:
        return float(self.x[1] - self.x[0])

    @property
    def spacingY(self) -> float:
        return float(self.y[1] - self.y[0])

    def _interpolate(self, x, y, derivative: str = '') -> np.ndarray:
        '''Cubic convolution at (x, y), or its partial derivative along `derivative` 'x' or 'y' (bathymetry.py:444-479).'''
        x, y = np.broadcast_arrays(np.asarray(x, dtype = float), np.asarray(y, dtype = float))
        nx, ny = self.x.size, self.y.size
        rawX = (x - self.x[0]) / self.spacingX
        rawY = (y - self.y[0]) / self.spacingY
        fx, fy = np.clip(rawX, 0.0, nx - 1.0), np.clip(rawY, 0.0, ny - 1.0)
        i, j = np.floor(fx).astype(np.intp), np.floor(fy).astype(np.intp)
        u, v = fx - i, fy - j
        wx = convolutionSlopeWeights(u) / self.spacingX if derivative == 'x' else convolutionWeights(u)
        wy = convolutionSlopeWeights(v) / self.spacingY if derivative == 'y' else convolutionWeights(v)
        result = np.zeros(x.shape)
        for a in range(4):
            row = np.clip(j - 1 + a, 0, ny - 1)
            along = np.zeros(x.shape)
            for b in range(4):
                along += wx[b] * self.z[row, np.clip(i - 1 + b, 0, nx - 1)]
            result += wy[a] * along
        if derivative == 'x':
            result[(rawX < 0.0) | (rawX > nx - 1.0)] = 0.0
        elif derivative == 'y':
            result[(rawY < 0.0) | (rawY > ny - 1.0)] = 0.0
        return result

    def elevation(self, x, y) -> np.ndarray:
        return self._interpolate(x, y)

    def gradient(self, x, y) -> np.ndarray:
        '''(dz_b/dx, dz_b/dy), shape (..., 2) (bathymetry.py:651-655).'''
        return np.stack([self._interpolate(x, y, 'x'), self._interpolate(x, y, 'y')], axis = -1)

    def normal(self, x, y) -> np.ndarray:
        '''Unit normal up out of the bed, shape (..., 3) (bathymetry.py:657-664).'''
        gradient = self.gradient(x, y)
        normal = np.concatenate([-gradient, np.ones(gradient.shape[:-1] + (1,))], axis = -1)
        return normal / np.linalg.norm(normal, axis = -1, keepdims = True)

    def depth(self, x, y) -> np.ndarray:
        return np.maximum(0.0, self.stillWaterLevel - self.elevation(x, y))

    def lowestElevation(self) -> float:
        '''The lowest grid elevation inside the bounds (bathymetry.py:709-717).'''
        startX, endX, startY, endY = self.bounds
        inX = (self.x >= startX - 1e-9) & (self.x <= endX + 1e-9)
        inY = (self.y >= startY - 1e-9) & (self.y <= endY + 1e-9)
        return float(np.min(self.z[np.ix_(inY, inX)]))

    def section(self, y: float) -> Bathymetry:
        '''The bed along x at fixed y over the bounds, through the grid's x positions and the bounds' ends (bathymetry.py:719-731).'''
        startX, endX = self.bounds[0], self.bounds[1]
        inside = self.x[(self.x > startX + 1e-9) & (self.x < endX - 1e-9)]
        x = np.concatenate([[startX], inside, [endX]])
        return Bathymetry(x, self.elevation(x, np.full_like(x, float(y))), self.stillWaterLevel)
