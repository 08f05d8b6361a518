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
