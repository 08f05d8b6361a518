'''

Solid boundaries as fixed particles, after pySurf's `sphBoundaries` (src/pysurf/sphBoundaries.py):
the geometry of the particle layers the pair sums see, and the description of the clamp the device
step applies to the fluid (the kernels enforce it; see `pair_sums.boundaryOf`).

'''

from __future__ import annotations

import numpy as np

from pysurf_amd.bathymetry import Bathymetry


class ContainerBoundary:

    '''An axis-aligned container, open at the top unless asked otherwise (sphBoundaries.py:61-147).'''

    def __init__(self, containerMin, containerMax, spacing: float, nLayers: int = 3, openTop: bool = True) -> None:
        self.containerMin = np.asarray(containerMin, dtype = float).copy()
        self.containerMax = np.asarray(containerMax, dtype = float).copy()
        self.spacing = spacing
        self.nLayers = nLayers
        self.openTop = openTop

    @property
    def dimensions(self) -> int:
        return len(self.containerMin)

    def positions(self) -> np.ndarray:
        '''The shell nLayers deep around the interior, less the top face when open (sphBoundaries.py:95-125).'''
        s, layers = self.spacing, self.nLayers
        offsets = (np.arange(layers) + 0.5) * s
        axes = []
        for axis, (low, high) in enumerate(zip(self.containerMin, self.containerMax)):
            pieces = [low - offsets[::-1], np.arange(low + s / 2.0, high, s)]
            if not (self.openTop and axis == self.dimensions - 1):
                pieces.append(high + offsets)
            axes.append(np.concatenate(pieces))
        grids = np.meshgrid(*axes, indexing = 'xy')
        points = np.column_stack([grid.ravel() for grid in grids])
        interior = np.all((points > self.containerMin) & (points < self.containerMax), axis = 1)
        if self.openTop:
            between = np.all((points[:, :-1] > self.containerMin[:-1]) & (points[:, :-1] < self.containerMax[:-1]),
                             axis = 1)
            interior = between & (points[:, -1] > self.containerMin[-1])
        return points[~interior]


class BedBoundary:

    '''

    A bed of any profile with optional offshore and shoreward walls; 2D, or extruded across a
    channel in 3D (sphBoundaries.py:149-344). Bed particles sit in layers offset along the bed
    normal by (layer + 1/2) spacing, one spacing apart along it.

    '''

    def __init__(self, bathymetry: Bathymetry, spacing: float, nLayers: int = 3, wallTop: float | None = None,
                 width: float | None = None, endWallTop: float | None = None, periodic: bool = False) -> None:
        self.bathymetry = bathymetry
        self.spacing = spacing
        self.nLayers = nLayers
        self.wallTop = wallTop
        self.width = width
        self.endWallTop = endWallTop
        self.periodic = periodic and width is not None
        self.wallX = bathymetry.startX
        self._pistonMask: np.ndarray | None = None

    def _profilePositions(self) -> np.ndarray:
        '''The 2D bed and wall layers in (x, z) (sphBoundaries.py:201-226).'''
        s = self.spacing
        bedX, bedZ = self.bathymetry.sample(s)
        bedNormal = self.bathymetry.normal(bedX)
        layers = [np.column_stack([bedX, bedZ]) - (layer + 0.5) * s * bedNormal for layer in range(self.nLayers)]
        for wallX, top, side in ((self.bathymetry.startX, self.wallTop, -1.0),
                                 (self.bathymetry.endX, self.endWallTop, 1.0)):
            if top is None:
                continue
            bottom = float(self.bathymetry.elevation(wallX)) - self.nLayers * s
            heights = np.arange(bottom + s / 2.0, top, s)
            for layer in range(self.nLayers):
                layers.append(np.column_stack([np.full_like(heights, wallX + side * (layer + 0.5) * s), heights]))
        return np.vstack(layers)

    def positions(self) -> np.ndarray:
        '''Boundary particle positions, (N, 2), or (N, 3) across a channel (sphBoundaries.py:228-258).'''
        profile = self._profilePositions()
        if self.width is None:
            return profile
        s, layers = self.spacing, self.nLayers
        offsets = (np.arange(layers) + 0.5) * s
        if self.periodic:
            across = np.arange(s / 2.0, self.width, s)
            return np.column_stack([np.repeat(profile[:, 0], across.size), np.tile(across, len(profile)),
                                    np.repeat(profile[:, 1], across.size)])
        across = np.concatenate([-offsets[::-1], np.arange(s / 2.0, self.width, s), self.width + offsets])
        extruded = np.column_stack([np.repeat(profile[:, 0], across.size), np.tile(across, len(profile)),
                                    np.repeat(profile[:, 1], across.size)])
        top = self.wallTop if self.wallTop is not None else float(np.max(self.bathymetry.z)) + layers * s
        bedX, _ = self.bathymetry.sample(s)
        heights = np.arange(float(np.min(self.bathymetry.z)) - layers * s + s / 2.0, top, s)
        xGrid, zGrid = np.meshgrid(bedX, heights, indexing = 'xy')
        keep = zGrid > self.bathymetry.elevation(xGrid) - layers * s
        xWall, zWall = xGrid[keep], zGrid[keep]
        sides = [np.column_stack([xWall, np.full_like(xWall, y), zWall])
                 for y in np.concatenate([-offsets, self.width + offsets])]
        return np.vstack([extruded] + sides)

    def attachPiston(self, particles) -> None:
        '''The offshore wall's particles as a piston (sphBoundaries.py:260-275).'''
        if self.wallTop is None:
            raise ValueError('A piston needs the offshore wall; give the boundary a wallTop.')
        self._pistonMask = (~particles.isFluid) & (particles.positions[:, 0] < self.bathymetry.startX - 0.25 * self.spacing)

    def movePiston(self, particles, x: float, velocity: float) -> None:
        '''Move the offshore wall to x at `velocity` (sphBoundaries.py:277-293).'''
        if self._pistonMask is None:
            raise RuntimeError('No piston attached. Call attachPiston first.')
        particles.positions[self._pistonMask, 0] += x - self.wallX
        particles.velocities[self._pistonMask] = 0.0
        particles.velocities[self._pistonMask, 0] = velocity
        self.wallX = x

    def description(self, pistonSpeed: float = 0.0) -> dict:
        '''The clamp as the device applies it (sphGpu.py:2141-2158), as a plain dict.'''
        bed, clearance = self.bathymetry, 0.25 * self.spacing
        return {'kind': 'bed', 'clearance': clearance, 'bedX': bed.x.copy(), 'bedZ': bed.z.copy(),
                'sideWalls': self.width is not None and not self.periodic, 'width': float(self.width or 0.0),
                'offshoreWall': self.wallTop is not None, 'wallX': float(self.wallX), 'pistonSpeed': pistonSpeed,
                'endLimit': bed.endX - clearance if self.endWallTop is not None else bed.endX}
