'''

The beds the boundary and the starts are laid on: a 2D profile, piecewise linear through vertices
(the behaviour of pySurf's `Bathymetry`, src/pysurf/bathymetry.py:66-308), and a 3D bed sampled on
a regular grid and read between grid points by cubic convolution (its `HeightField`,
bathymetry.py:370-850), the parts the delta-SPH runs use.

'''

from __future__ import annotations

import math

import numpy as np

# Keys' cubic convolution kernel with a = -1/2 (Keys 1981): the weight of each of the four grid
# points around a query a fraction u past the second, as the coefficients of u^3, u^2, u, 1
KEYS = ((-0.5, 1.0, -0.5, 0.0),
        (1.5, -2.5, 0.0, 1.0),
        (-1.5, 2.0, 0.5, 0.0),
        (0.5, -0.5, 0.0, 0.0))

# A grid point within this of a bound counts as on it
ON_BOUND = 1e-9


class Bathymetry:

    '''

    A 2D bed z_b(x), linear between vertices (x, z) with x strictly increasing and level beyond the
    ends, under still water at `stillWaterLevel` (bathymetry.py:66-113).

    '''

    def __init__(self, x, z, stillWaterLevel: float) -> None:
        vertices = np.asarray(x, dtype = float), np.asarray(z, dtype = float)
        if vertices[0].ndim != 1 or vertices[0].size < 2 or vertices[1].shape != vertices[0].shape:
            raise ValueError('A bed profile needs at least two (x, z) vertices of matching length.')
        if not np.isfinite(vertices).all():
            raise ValueError('Bed vertices must be finite.')
        if (vertices[0][1:] <= vertices[0][:-1]).any():
            raise ValueError('Bed vertex x must increase strictly, so the bed is a function of x.')
        self.x, self.z = vertices
        self.stillWaterLevel = float(stillWaterLevel)
        # each segment's dz/dx, the last one's holding beyond the end
        self._gradients = np.diff(self.z) / np.diff(self.x)

    @classmethod
    def planeBeach(cls, depth: float, toeX: float, slope: float, endX: float, startX: float = 0.0) -> Bathymetry:
        '''Level at z = 0 from startX to the toe, rising at `slope` to endX, still water `depth` over the level part (bathymetry.py:115-144).'''
        if not startX < toeX < endX:
            raise ValueError('A plane beach needs startX < toeX < endX.')
        if min(slope, depth) <= 0.0:
            raise ValueError('Depth and slope must be positive.')
        return cls([startX, toeX, endX], [0.0, 0.0, (endX - toeX) * slope], stillWaterLevel = depth)

    @property
    def startX(self) -> float:
        return float(self.x[0])

    @property
    def endX(self) -> float:
        return float(self.x[-1])

    def elevation(self, x) -> np.ndarray:
        '''z_b(x) (bathymetry.py:221-225).'''
        return np.interp(x, self.x, self.z)

    def depth(self, x) -> np.ndarray:
        '''The still water's depth, never below zero (bathymetry.py:227-231).'''
        return np.maximum(0.0, self.stillWaterLevel - self.elevation(x))

    def slope(self, x) -> np.ndarray:
        '''dz_b/dx of the segment x lies on; a vertex belongs to the segment it starts (bathymetry.py:233-240).'''
        segment = np.searchsorted(self.x, x, side = 'right') - 1
        return self._gradients[np.clip(segment, 0, self._gradients.size - 1)]

    def normal(self, x) -> np.ndarray:
        '''The unit normal pointing up out of the bed, (-dz_b/dx, 1) normalised, shape (..., 2) (bathymetry.py:242-250).'''
        gradient = np.asarray(self.slope(x), dtype = float)
        length = np.sqrt(1.0 + gradient**2)
        return np.stack([-gradient / length, 1.0 / length], axis = -1)

    def sample(self, spacing: float) -> tuple[np.ndarray, np.ndarray]:
        '''(x, z) of points evenly spaced along the bed's length, about `spacing` apart, both ends included (bathymetry.py:272-294).'''
        along = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(self.x), np.diff(self.z)))])
        stations = np.linspace(0.0, along[-1], max(2, int(round(along[-1] / spacing)) + 1))
        return np.interp(stations, along, self.x), np.interp(stations, along, self.z)


def _cubic(table, u: np.ndarray, u2: np.ndarray, u3: np.ndarray) -> np.ndarray:
    '''Each row (c3, c2, c1, c0) of `table` evaluated as c3 u^3 + c2 u^2 + c1 u + c0, stacked, shape (rows, ...).'''
    return np.stack([c3 * u3 + c2 * u2 + c1 * u + c0 for c3, c2, c1, c0 in table])


def convolutionWeights(u) -> np.ndarray:
    '''Keys' weights of the four grid points around a fraction u, shape (4, ...); they sum to one and reproduce quadratics (bathymetry.py:310-324).'''
    u = np.asarray(u, dtype = float)
    u2 = u * u
    return _cubic(KEYS, u, u2, u2 * u)


def convolutionSlopeWeights(u) -> np.ndarray:
    '''The u-derivatives of convolutionWeights, shape (4, ...) (bathymetry.py:326-333).'''
    u = np.asarray(u, dtype = float)
    slopes = [(0.0, 3.0 * c3, 2.0 * c2, c1) for c3, c2, c1, _ in KEYS]
    return _cubic(slopes, u, u * u, 0.0)


def _accumulate(weights, terms) -> np.ndarray:
    '''sum_k weights[k] terms[k], added to zero in order of k.'''
    total = 0.0
    for weight, term in zip(weights, terms):
        total = total + weight * term
    return total


class HeightField:

    '''

    A 3D bed z_b(x, y) sampled on a regular grid (`z` shaped (len(y), len(x))), x along the waves,
    y across them, z up (bathymetry.py:370-442). Between grid points it is read by cubic
    convolution over the 4 x 4 points around a query, so the surface and its gradient are
    continuous and quadratics come out exact; beyond the grid a query takes the edge's value, with
    no slope across the edge. `bounds` (startX, endX, startY, endY) is the region the bed covers,
    the tank's walls, inside the grid.

    '''

    def __init__(self, x, y, z, stillWaterLevel: float, bounds: tuple | None = None) -> None:
        self.x = np.asarray(x, dtype = float)
        self.y = np.asarray(y, dtype = float)
        self.z = np.asarray(z, dtype = float)
        self.stillWaterLevel = float(stillWaterLevel)
        if any(axis.ndim != 1 or axis.size < 4 for axis in (self.x, self.y)):
            raise ValueError('A height field needs at least four grid points along x and y.')
        if self.z.shape != (self.y.size, self.x.size):
            raise ValueError('The elevations must be shaped (len(y), len(x)).')
        if not all(np.isfinite(values).all() for values in (self.x, self.y, self.z)):
            raise ValueError('Height field values must be finite.')
        for axis in (self.x, self.y):
            steps = np.diff(axis)
            if (steps <= 0.0).any() or not np.allclose(steps, steps[0], rtol = 1e-9, atol = 0.0):
                raise ValueError('The grid must increase in equal steps along each axis.')
        self.bounds = tuple(float(value) for value in (bounds if bounds is not None
                                                        else (self.x[0], self.x[-1], self.y[0], self.y[-1])))
        startX, endX, startY, endY = self.bounds
        fits = [axis[0] - ON_BOUND <= low < high <= axis[-1] + ON_BOUND
                for axis, low, high in ((self.x, startX, endX), (self.y, startY, endY))]
        if not all(fits):
            raise ValueError('The bounds must lie inside the grid.')

    @classmethod
    def fromFunction(cls, function, bounds: tuple, gridSpacing: float, stillWaterLevel: float,
                     padding: int = 2) -> HeightField:

        '''

        The bed `function`(x, y) sampled over `bounds` and `padding` grid steps beyond them on every
        side (bathymetry.py:481-516). Each axis takes the step nearest `gridSpacing` that divides its
        bounds into whole cells.

        '''

        startX, endX, startY, endY = map(float, bounds)
        if not (endX > startX and endY > startY and gridSpacing > 0.0):
            raise ValueError('A height field needs positive extents and grid spacing.')

        def axis(low, high):
            cells = max(1, int(math.ceil((high - low) / gridSpacing - 1e-9)))
            return low + (high - low) / cells * np.arange(-padding, cells + padding + 1)

        gridX, gridY = axis(startX, endX), axis(startY, endY)
        table = np.asarray(function(*np.meshgrid(gridX, gridY, indexing = 'xy')), dtype = float)
        return cls(gridX, gridY, table, stillWaterLevel, bounds = (startX, endX, startY, endY))

    @classmethod
    def sechRidge(cls, k: float, depth: float = 1.0, toeX: float = 5.225, slope: float = 1.0 / 15.0,
                  endX: float = 19.0, halfWidth: float = 4.0, startX: float = 0.0,
                  gridSpacing: float | None = None) -> HeightField:
        '''The ridge z_b / d = slope max(x / d - toeX, 0) sech^2(k y / d), in a tank [startX, endX] x [-halfWidth, halfWidth] in units of d (bathymetry.py:518-553).'''

        def bed(x, y):
            return depth * slope * np.maximum(x / depth - toeX, 0.0) / np.cosh(k * y / depth)**2

        return cls.fromFunction(bed, (startX * depth, endX * depth, -halfWidth * depth, halfWidth * depth),
                                depth / 64.0 if gridSpacing is None else gridSpacing, stillWaterLevel = depth)

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
        cosine, sine = math.cos(turn), math.sin(turn)

        def bed(x, y):
            offshore, across = x / depth - toeX, y / depth
            along = toeX + offshore * cosine + across * sine
            off = -offshore * sine + across * cosine
            crest = np.minimum(slope * np.maximum(along - toeX, 0.0), height)
            return depth * crest * np.exp(-off**2 / (2.0 * width**2))

        return cls.fromFunction(bed, (0.0, length * depth, -halfWidth * depth, halfWidth * depth),
                                depth / 64.0 if gridSpacing is None else gridSpacing, stillWaterLevel = depth)

    @property
    def spacingX(self) -> float:
        return float(self.x[1] - self.x[0])

    @property
    def spacingY(self) -> float:
        return float(self.y[1] - self.y[0])

    def _stencil(self, coordinate: np.ndarray, axis: np.ndarray, step: float, slope: bool):
        '''Along one axis: the four grid indices around each query (clamped to the grid), their weights (of the value, or with `slope` of its derivative), and whether the query lies beyond the grid.'''
        last = axis.size - 1
        position = (coordinate - axis[0]) / step
        clamped = np.clip(position, 0.0, float(last))
        cell = np.floor(clamped).astype(np.intp)
        fraction = clamped - cell
        weights = convolutionSlopeWeights(fraction) / step if slope else convolutionWeights(fraction)
        indices = [np.clip(cell + shift, 0, last) for shift in (-1, 0, 1, 2)]
        return indices, weights, (position < 0.0) | (position > last)

    def _interpolate(self, x, y, derivative: str = '') -> np.ndarray:
        '''The convolution at (x, y): the elevation, or its derivative along `derivative` ('x' or 'y') (bathymetry.py:444-479).'''
        x, y = np.broadcast_arrays(np.asarray(x, dtype = float), np.asarray(y, dtype = float))
        columns, wx, outsideX = self._stencil(x, self.x, self.spacingX, derivative == 'x')
        rows, wy, outsideY = self._stencil(y, self.y, self.spacingY, derivative == 'y')
        value = _accumulate(wy, [_accumulate(wx, [self.z[row, column] for column in columns]) for row in rows])
        if derivative:
            value = np.where(outsideX if derivative == 'x' else outsideY, 0.0, value)
        return value

    def elevation(self, x, y) -> np.ndarray:
        '''z_b(x, y) (bathymetry.py:645-649).'''
        return self._interpolate(x, y)

    def gradient(self, x, y) -> np.ndarray:
        '''(dz_b/dx, dz_b/dy), shape (..., 2) (bathymetry.py:651-655).'''
        return np.stack([self._interpolate(x, y, 'x'), self._interpolate(x, y, 'y')], axis = -1)

    def normal(self, x, y) -> np.ndarray:
        '''The unit normal pointing up out of the bed, (-grad z_b, 1) normalised, shape (..., 3) (bathymetry.py:657-664).'''
        slope = self.gradient(x, y)
        up = np.concatenate([-slope, np.ones(slope.shape[:-1] + (1,))], axis = -1)
        return up / np.linalg.norm(up, axis = -1, keepdims = True)

    def depth(self, x, y) -> np.ndarray:
        '''The still water's depth, never below zero (bathymetry.py:666-670).'''
        return np.maximum(0.0, self.stillWaterLevel - self.elevation(x, y))

    def _insideBounds(self):
        '''Masks of the grid's x and y positions within the bounds.'''
        startX, endX, startY, endY = self.bounds
        return ((self.x >= startX - ON_BOUND) & (self.x <= endX + ON_BOUND),
                (self.y >= startY - ON_BOUND) & (self.y <= endY + ON_BOUND))

    def lowestElevation(self) -> float:
        '''The lowest elevation at a grid point within the bounds (bathymetry.py:709-717).'''
        inX, inY = self._insideBounds()
        return float(self.z[inY][:, inX].min())

    def section(self, y: float) -> Bathymetry:
        '''The 2D bed along x at fixed y across the bounds, with vertices at both bounds and the grid's x strictly between (bathymetry.py:719-731).'''
        startX, endX = self.bounds[:2]
        between = self.x[(self.x > startX + ON_BOUND) & (self.x < endX - ON_BOUND)]
        x = np.r_[startX, between, endX]
        return Bathymetry(x, self.elevation(x, np.full(x.shape, float(y))), self.stillWaterLevel)
