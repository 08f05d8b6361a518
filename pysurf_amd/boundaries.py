'''

Solid boundaries as fixed particles, after pySurf's `sphBoundaries` (src/pysurf/sphBoundaries.py):
the geometry of the particle layers the pair sums see, and the description of the clamp the device
step applies to the fluid (the kernels enforce it; see `pair_sums.boundaryOf`).

'''

from __future__ import annotations

import numpy as np

from pysurf_amd.bathymetry import Bathymetry, HeightField


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


class HeightFieldBoundary:

    '''

    A height-field bed in a rectangular tank whose walls stand on the field's bounds: offshore at
    x = startX, shoreward at x = endX and at y = startY and y = endY (sphBoundaries.py:498-640). Bed
    particles sit on the lattice's columns, (i + 1/2) s from the walls, in layers below the bed along
    its local normal at (layer + 1/2) s; each wall's layers stand (layer + 1/2) s behind its plane on
    the fluid lattice's heights, from below the bed's deepest layer up to `wallTop`, the side walls
    running past the end walls to close the corners.

    '''

    def __init__(self, field: HeightField, spacing: float, nLayers: int = 3, wallTop: float = 1.0,
                 offshoreWall: bool = True, endWall: bool = True) -> None:
        self.field = field
        self.spacing = spacing
        self.nLayers = nLayers
        self.wallTop = wallTop
        self.offshoreWall = offshoreWall
        self.endWall = endWall

    @property
    def dimensions(self) -> int:
        return 3

    @property
    def latticeFloor(self) -> float:
        '''The level the fluid lattice's heights count from: the bed's lowest point (sphBoundaries.py:545-548).'''
        return self.field.lowestElevation()

    def columns(self) -> tuple[np.ndarray, np.ndarray]:
        '''The lattice's column positions along x and y inside the walls (sphBoundaries.py:550-556).'''
        s = self.spacing
        startX, endX, startY, endY = self.field.bounds
        return np.arange(startX + 0.5 * s, endX, s), np.arange(startY + 0.5 * s, endY, s)

    def positions(self) -> np.ndarray:
        '''Boundary particle positions, (N, 3): the bed's layers, then the walls (sphBoundaries.py:558-595).'''
        s, layers, field = self.spacing, self.nLayers, self.field
        startX, endX, startY, endY = field.bounds
        xs, ys = self.columns()
        gridX, gridY = np.meshgrid(xs, ys, indexing = 'xy')
        gridX, gridY = gridX.ravel(), gridY.ravel()
        surface = np.column_stack([gridX, gridY, field.elevation(gridX, gridY)])
        normal = field.normal(gridX, gridY)
        blocks = [surface - (layer + 0.5) * s * normal for layer in range(layers)]
        heights = np.arange(self.latticeFloor - (layers + 1) * s + 0.5 * s, self.wallTop, s)

        def wall(lineX, lineY):
            bottom = field.elevation(np.clip(lineX, startX, endX), np.clip(lineY, startY, endY)) - layers * s
            x, y = np.repeat(lineX, heights.size), np.repeat(lineY, heights.size)
            z = np.tile(heights, lineX.size)
            keep = z > np.repeat(bottom, heights.size)
            return np.column_stack([x[keep], y[keep], z[keep]])

        offsets = (np.arange(layers) + 0.5) * s
        for offset in offsets:
            if self.offshoreWall:
                blocks.append(wall(np.full(ys.size, startX - offset), ys))
            if self.endWall:
                blocks.append(wall(np.full(ys.size, endX + offset), ys))
        alongAll = np.concatenate([startX - offsets[::-1], xs, endX + offsets])
        for offset in offsets:
            blocks.append(wall(alongAll, np.full(alongAll.size, startY - offset)))
            blocks.append(wall(alongAll, np.full(alongAll.size, endY + offset)))
        return np.vstack(blocks)

    def enforce(self, particles) -> None:

        '''

        Return fluid that crossed the bed or a wall a quarter spacing clear of it, removing its
        velocity into the solid: below the bed a particle rises vertically and loses its velocity
        into the bed along the bed's normal there (sphBoundaries.py:597-640).

        '''

        fluid = np.nonzero(particles.isFluid)[0]
        positions, velocities = particles.positions, particles.velocities
        clearance = 0.25 * self.spacing
        field = self.field
        startX, endX, startY, endY = field.bounds
        bed = field.elevation(positions[fluid, 0], positions[fluid, 1])
        below = positions[fluid, 2] < bed + clearance
        under = fluid[below]
        if under.size:
            positions[under, 2] = bed[below] + clearance
            normal = field.normal(positions[under, 0], positions[under, 1])
            inward = np.sum(velocities[under] * normal, axis = 1)
            into = inward < 0.0
            velocities[under[into]] -= inward[into, None] * normal[into]
        low = fluid[positions[fluid, 1] < startY + clearance]
        positions[low, 1] = startY + clearance
        velocities[low, 1] = np.maximum(velocities[low, 1], 0.0)
        high = fluid[positions[fluid, 1] > endY - clearance]
        positions[high, 1] = endY - clearance
        velocities[high, 1] = np.minimum(velocities[high, 1], 0.0)
        if self.offshoreWall:
            behind = fluid[positions[fluid, 0] < startX + clearance]
            positions[behind, 0] = startX + clearance
            velocities[behind, 0] = np.maximum(velocities[behind, 0], 0.0)
        limit = endX - clearance if self.endWall else endX
        beyond = fluid[positions[fluid, 0] > limit]
        positions[beyond, 0] = limit
        velocities[beyond, 0] = np.minimum(velocities[beyond, 0], 0.0)

    def description(self) -> dict:
        '''The clamp as the device applies it (sphGpu.py:2159-2171), as a plain dict.'''
        field, clearance = self.field, 0.25 * self.spacing
        startX, endX, startY, endY = field.bounds
        return {'kind': 'heightField', 'clearance': clearance, 'table': field.z.copy(), 'originX': float(field.x[0]),
                'originY': float(field.y[0]), 'stepX': field.spacingX, 'stepY': field.spacingY,
                'lowY': startY + clearance, 'highY': endY - clearance, 'offshoreWall': self.offshoreWall,
                'lowX': startX + clearance, 'highX': endX - clearance if self.endWall else endX}
