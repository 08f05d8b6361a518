'''

The breaking read-out of a 2D particle cloud, after pySurf's `waveBreaking`
(src/pysurf/waveBreaking.py:129-657, 658-906): the free-surface profile, the trough ahead of the
crest, the waterline, the crest's front face read one of two ways, and the tracker that records the
breaking point from them (Grilli et al. 1997: breaking where the front face first turns vertical).

- 'cells': the lean of the face in cells one spacing square; breaking at the first of `persistence`
  consecutive updates leaning forward of vertical by a spacing.
- 'contour': the one-half contour of the kernel-smoothed fluid fraction, sorted into the surface,
  pockets of air and drops of water; the face's least dx/ds along its arc; breaking at the start of
  the last run of updates with the face vertical before the jet's touchdown (a pocket of air closing
  under it), or of the longest such run without one.

A particle stands for the cell of water around it, so elevations include the half spacing above
the topmost particle centres, as the reference's do.

'''

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def freeSurfaceProfile(positions: np.ndarray, x0: float, x1: float, columnWidth: float) -> tuple[np.ndarray, np.ndarray]:
    '''Column centres and the highest fluid particle plus half a column in each column, NaN where empty (waveBreaking.py:129-162).'''
    count = max(1, int(np.ceil((x1 - x0) / columnWidth)))
    centers = x0 + (np.arange(count) + 0.5) * columnWidth
    elevation = np.full(count, np.nan)
    positions = np.asarray(positions, dtype = float)
    if len(positions) == 0:
        return centers, elevation
    columns = np.floor((positions[:, 0] - x0) / columnWidth).astype(np.int64)
    keep = (columns >= 0) & (columns < count)
    if np.any(keep):
        top = np.full(count, -np.inf)
        np.maximum.at(top, columns[keep], positions[keep, 1])
        elevation = np.where(np.isfinite(top), top + 0.5 * columnWidth, np.nan)
    return centers, elevation


def troughAhead(centers: np.ndarray, elevation: np.ndarray, crestX: float, reach: float = np.inf) -> float:

    '''

    The lowest free surface over the wet columns from the crest to the first empty column or to
    `reach` past it, after a three-column running median with reflected ends; NaN when the column
    just ahead of the crest is empty (waveBreaking.py:164-204).

    '''

    centers = np.asarray(centers)
    ahead = np.flatnonzero((centers > crestX) & (centers <= crestX + reach))
    wet = np.isfinite(np.asarray(elevation)[ahead])
    run = ahead[:int(np.argmin(wet)) if not wet.all() else wet.size]
    if run.size == 0:
        return float('nan')
    levels = np.asarray(elevation, dtype = float)[run]
    if levels.size < 3:
        return float(np.max(levels))
    padded = np.concatenate([levels[1:2], levels, levels[-2:-1]])
    smoothed = np.median(np.stack([padded[:-2], padded[1:-1], padded[2:]]), axis = 0)
    return float(np.min(smoothed))


def frontFaceLean(positions: np.ndarray, crestX: float, stillWaterLevel: float, spacing: float, reach: float,
                  shoreLimit: float = np.inf) -> float:

    '''

    How far the crest's shoreward face leans forward of vertical (waveBreaking.py:559-634).

    The water from half a spacing behind the crest to `reach` ahead of it and more than half a
    spacing above still water is binned into cells one spacing square, each particle marking the 2 x 2
    cells its own cell overlaps. In each band the face is the shoreward end of the run of marked
    cells that starts at the crest column (bands whose run reaches the window's end are left out);
    the lean is the largest amount a band's face stands shoreward of the face of any band below it.
    Returns -inf when fewer than two bands have a face.

    '''

    occupied, windowStart = faceCells(positions, crestX, stillWaterLevel, spacing, reach, shoreLimit)
    return leanOfCells(occupied, windowStart, spacing)


def faceWindow(crestX: float, stillWaterLevel: float, spacing: float, reach: float,
               shoreLimit: float = np.inf) -> tuple[float, float, float, int]:
    '''The face window's start, end, band base and column count (waveBreaking.py:575-590).'''
    windowStart = crestX - 0.5 * spacing
    windowEnd = min(crestX + reach, shoreLimit)
    base = stillWaterLevel + 0.5 * spacing
    nColumns = max(1, int(np.ceil((windowEnd - windowStart) / spacing)))
    return windowStart, windowEnd, base, nColumns


def faceCells(positions: np.ndarray, crestX: float, stillWaterLevel: float, spacing: float, reach: float,
              shoreLimit: float = np.inf) -> tuple[np.ndarray, float]:
    '''The marked cells of the face window, (bands, columns) upward from still water, and the window's start.'''
    windowStart, windowEnd, base, nColumns = faceWindow(crestX, stillWaterLevel, spacing, reach, shoreLimit)
    x, z = positions[:, 0], positions[:, 1]
    window = (x >= windowStart - 0.5 * spacing) & (x < windowEnd + 0.5 * spacing) & (z > base - 0.5 * spacing)
    if not np.any(window):
        return np.zeros((0, nColumns), dtype = bool), windowStart
    firstColumn = np.floor((x[window] - 0.5 * spacing - windowStart) / spacing).astype(np.int64)
    firstBand = np.floor((z[window] - 0.5 * spacing - base) / spacing).astype(np.int64)
    occupied = np.zeros((int(firstBand.max()) + 2, nColumns), dtype = bool)
    for columnStep in (0, 1):
        for bandStep in (0, 1):
            column, band = firstColumn + columnStep, firstBand + bandStep
            inside = (column >= 0) & (column < nColumns) & (band >= 0)
            occupied[band[inside], column[inside]] = True
    return occupied, windowStart


def leanOfCells(occupied: np.ndarray, windowStart: float, spacing: float) -> float:
    '''The lean of a (bands, columns) occupancy of the face window, bands upward from still water (waveBreaking.py:613-634).'''
    faces = []
    for row in occupied:
        start = np.flatnonzero(row[:2])
        if start.size == 0:
            continue
        rest = row[start[0]:]
        if np.all(rest):
            continue
        faces.append(windowStart + (start[0] + int(np.argmin(rest))) * spacing)
    if len(faces) < 2:
        return -np.inf
    faces = np.asarray(faces)
    lowestBelow = np.minimum.accumulate(faces)
    return float(np.max(faces[1:] - lowestBelow[:-1]))


def waterlineX(positions: np.ndarray, bathymetry, wetHeight: float, minimumParticles: int = 3) -> float:
    '''The `minimumParticles`-th most landward particle within `wetHeight` of the bed (waveBreaking.py:636-655).'''
    nearBed = positions[positions[:, 1] - bathymetry.elevation(positions[:, 0]) < wetHeight]
    if len(nearBed) < minimumParticles:
        return float('nan')
    return float(np.sort(nearBed[:, 0])[-minimumParticles])


def waterlineFromLandward(landward: np.ndarray, bathymetry, wetHeight: float, complete: bool,
                          minimumParticles: int = 3) -> float | None:

    '''

    waterlineX from the most landward candidates only: `landward` (m, 2) holds the particles of
    largest x among a superset of those within `wetHeight` of the bed, by descending x, and
    `complete` says whether it holds the whole superset. The exact test on them decides the answer
    whenever `minimumParticles` of them pass it or the candidates are all there are; otherwise
    returns None, and the full positions must decide.

    '''

    landward = landward[np.isfinite(landward[:, 0])]
    nearBed = landward[landward[:, 1] - bathymetry.elevation(landward[:, 0]) < wetHeight]
    if len(nearBed) >= minimumParticles:
        return float(np.sort(nearBed[:, 0])[-minimumParticles])
    return float('nan') if complete else None


# -- The smoothed surface ('contour') -- #

def cubicSpline2D(distance: np.ndarray, h: float) -> np.ndarray:
    '''The 2D cubic spline W(r, h), sigma = 10 / (7 pi h^2), support 2h (sphKernels.py:267-303).'''
    q = np.asarray(distance, dtype = float) / h
    sigma = 10.0 / (7.0 * math.pi * h * h)
    weight = np.zeros_like(q)
    inner = q < 1.0
    weight[inner] = sigma * (1.0 - 1.5 * q[inner]**2 + 0.75 * q[inner]**3)
    outer = (q >= 1.0) & (q < 2.0)
    weight[outer] = sigma * 0.25 * (2.0 - q[outer])**3
    return weight


def fractionGrid(spacing: float, lower, upper) -> tuple[np.ndarray, np.ndarray]:
    '''The fluid fraction's nodes: x and y half a spacing apart from `lower` to past `upper` (waveBreaking.py:240-243).'''
    step = 0.5 * spacing
    return np.arange(lower[0], upper[0] + step, step), np.arange(lower[1], upper[1] + step, step)


def fluidFraction(positions: np.ndarray, spacing: float, lower, upper, smoothingRatio: float = 1.3,
                  volumes: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray, np.ndarray]:

    '''

    The kernel-smoothed fluid fraction phi(x) = sum_j V_j W(|x - x_j|, h) on the nodes of
    `fractionGrid`, h = `smoothingRatio` spacings, V_j = `volumes` or s^2 (waveBreaking.py:206-262).
    phi is one inside the water and one half where a lattice ends, so its one-half contour is the
    free surface. Returns the nodes' x, y and phi, shape (ny, nx).

    Each particle of the window widened by 2h is binned by the node at or below it; a node sums
    the bins around it offset by offset, rows outer, and each bin's particles in order.

    '''

    step = 0.5 * spacing
    h = smoothingRatio * spacing
    gridX, gridY = fractionGrid(spacing, lower, upper)
    fraction = np.zeros((len(gridY), len(gridX)))
    positions = np.asarray(positions, dtype = float).reshape(-1, 2)
    reach = int(np.ceil(2.0 * h / step))
    inside = ((positions[:, 0] > lower[0] - 2.0 * h) & (positions[:, 0] < upper[0] + 2.0 * h)
              & (positions[:, 1] > lower[1] - 2.0 * h) & (positions[:, 1] < upper[1] + 2.0 * h))
    points = positions[inside]
    areas = np.full(len(points), spacing**2) if volumes is None else np.asarray(volumes, dtype = float)[inside]
    binX = np.floor((points[:, 0] - gridX[0]) / step).astype(np.int64)
    binY = np.floor((points[:, 1] - gridY[0]) / step).astype(np.int64)
    flat = fraction.ravel()
    for offsetY in range(-reach, reach + 2):
        for offsetX in range(-reach, reach + 2):
            nodeX, nodeY = binX + offsetX, binY + offsetY
            valid = (nodeX >= 0) & (nodeX < len(gridX)) & (nodeY >= 0) & (nodeY < len(gridY))
            distance = np.hypot(gridX[nodeX[valid]] - points[valid, 0], gridY[nodeY[valid]] - points[valid, 1])
            flat += np.bincount(nodeY[valid] * len(gridX) + nodeX[valid], minlength = flat.size,
                                weights = areas[valid] * cubicSpline2D(distance, h))
    return gridX, gridY, fraction


def halfContours(gridX: np.ndarray, gridY: np.ndarray, fraction: np.ndarray) -> list:
    '''The one-half contours of a gridded fraction as polylines, by contourpy (waveBreaking.py:300-308).'''
    import contourpy
    generator = contourpy.contour_generator(gridX, gridY, fraction, line_type = contourpy.LineType.Separate)
    return [np.asarray(line) for line in generator.lines(0.5)]


def fluidSurface(positions: np.ndarray, spacing: float, lower, upper, smoothingRatio: float = 1.3,
                 volumes: np.ndarray | None = None) -> list:
    '''The one-half contours of the smoothed fluid fraction (waveBreaking.py:264-298).'''
    return halfContours(*fluidFraction(positions, spacing, lower, upper, smoothingRatio, volumes))


@dataclass
class ParticleSurface:

    '''

    The free surface within a window (waveBreaking.py:310-340): `surface`, the longest contour that
    ends on the window's edges, ordered from its end of least x, (0, 2) when there is none;
    `pockets`, closed contours around air; `drops`, closed contours around water.

    '''

    surface: np.ndarray
    pockets: list
    drops: list

    @staticmethod
    def area(loop: np.ndarray) -> float:
        '''The area a closed contour encloses (the shoelace formula).'''
        x, y = loop[:, 0], loop[:, 1]
        return 0.5 * abs(float(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1))))


def sortContours(gridX: np.ndarray, gridY: np.ndarray, fraction: np.ndarray) -> ParticleSurface:

    '''

    The one-half contours of a gridded fraction sorted into the surface, pockets and drops
    (waveBreaking.py:342-402): an open contour is part of the surface the window cuts and the
    longest is kept; a closed one holds air when the mean fraction of the nodes inside it is below
    one half (a loop too small to hold a node: the node nearest its vertex mean decides).

    '''

    from matplotlib.path import Path
    surface, pockets, drops = np.zeros((0, 2)), [], []
    for line in halfContours(gridX, gridY, fraction):
        if len(line) < 3 or not np.allclose(line[0], line[-1]):
            if len(line) > len(surface):
                surface = line if line[0, 0] <= line[-1, 0] else line[::-1]
            continue
        columns = np.flatnonzero((gridX >= line[:, 0].min()) & (gridX <= line[:, 0].max()))
        rows = np.flatnonzero((gridY >= line[:, 1].min()) & (gridY <= line[:, 1].max()))
        nodeX, nodeY = np.meshgrid(gridX[columns], gridY[rows])
        inside = Path(line).contains_points(np.column_stack([nodeX.ravel(), nodeY.ravel()]))
        if inside.any():
            enclosed = float(np.mean(fraction[np.ix_(rows, columns)].ravel()[inside]))
        else:
            centre = line.mean(axis = 0)
            enclosed = float(fraction[int(np.argmin(np.abs(gridY - centre[1]))), int(np.argmin(np.abs(gridX - centre[0])))])
        (pockets if enclosed < 0.5 else drops).append(line)
    return ParticleSurface(surface = surface, pockets = pockets, drops = drops)


def particleSurface(positions: np.ndarray, spacing: float, lower, upper, smoothingRatio: float = 1.3,
                    volumes: np.ndarray | None = None) -> ParticleSurface:
    '''The window's surface, pockets and drops from the particles (waveBreaking.py:342-402).'''
    return sortContours(*fluidFraction(positions, spacing, lower, upper, smoothingRatio, volumes))


def airShare(loop: np.ndarray, surface: np.ndarray, spacing: float) -> float:

    '''

    The share of a closed contour's area on the air side of an open surface ordered by x: outside
    the polygon the surface makes when closed through the water below everything
    (waveBreaking.py:404-443). The loop's inside is sampled a quarter spacing apart, or at its vertex
    mean when too small. NaN when the surface has fewer than two points.

    '''

    from matplotlib.path import Path
    if len(surface) < 2:
        return float('nan')
    floor = min(float(surface[:, 1].min()), float(loop[:, 1].min())) - spacing
    water = Path(np.vstack([surface, [[surface[-1, 0], floor], [surface[0, 0], floor]]]))
    step = 0.25 * spacing
    nodeX, nodeY = np.meshgrid(np.arange(loop[:, 0].min(), loop[:, 0].max() + step, step),
                               np.arange(loop[:, 1].min(), loop[:, 1].max() + step, step))
    nodes = np.column_stack([nodeX.ravel(), nodeY.ravel()])
    nodes = nodes[Path(loop).contains_points(nodes)]
    if len(nodes) == 0:
        nodes = loop.mean(axis = 0, keepdims = True)
    return float(np.mean(~water.contains_points(nodes)))


@dataclass
class ContourFace:

    '''

    The crest and front face read from the smoothed surface (waveBreaking.py:445-469): the crest's x
    and height above still water (NaN without a surface), the least dx/ds on the face and where it
    lies (NaN when the face is too short), and the window's ParticleSurface (None without a window).

    '''

    crestX: float
    crestHeight: float
    slope: float
    faceX: float
    window: ParticleSurface | None


def contourWindow(spacing: float, crestX: float, bathymetry, reach: float, windowDepth: float):

    '''

    The window the face is read in (waveBreaking.py:520-530): from `reach` behind the crest to twice
    `reach` ahead, stopping short of where the bed comes within three spacings of the lower edge;
    from `windowDepth` below still water (in shallow water half of the depth under the crest left
    after three spacings) to twice `reach` above it. Returns (lower, upper), or None when less than
    a spacing of depth or two spacings of length remain.

    '''

    level = bathymetry.stillWaterLevel
    depth = min(windowDepth, 0.5 * (float(bathymetry.depth(crestX)) - 3.0 * spacing))
    if depth < spacing:
        return None
    bottom = level - depth
    ahead = np.linspace(crestX, min(crestX + 2.0 * reach, bathymetry.endX), 400)
    shallow = np.flatnonzero(bathymetry.elevation(ahead) > bottom - 3.0 * spacing)
    right = float(ahead[shallow[0]]) if shallow.size else float(ahead[-1])
    if right <= crestX + 2.0 * spacing:
        return None
    return np.array([crestX - reach, bottom]), np.array([right, level + 2.0 * reach])


def faceOfSurface(window: ParticleSurface, spacing: float, stillWaterLevel: float, arcLength: float = 2.0) -> ContourFace:

    '''

    The crest and the least dx/ds along the front face of a window's surface (waveBreaking.py:531-557):
    the face runs from the surface's highest point to its shoreward end, resampled every half
    spacing of arc; its slope is the change in x over `arcLength` spacings of arc, taken only where
    the face stands more than half a spacing above still water.

    '''

    missing = float('nan')
    surface = window.surface
    if len(surface) < 8:
        return ContourFace(missing, missing, missing, missing, window)
    top = int(np.argmax(surface[:, 1]))
    crest, height = float(surface[top, 0]), float(surface[top, 1] - stillWaterLevel)
    front = surface[top:]
    arc = np.concatenate([[0.0], np.cumsum(np.hypot(*np.diff(front, axis = 0).T))])
    if arc[-1] < max(4.0, arcLength + 2.0) * spacing:
        return ContourFace(crest, height, missing, missing, window)
    steps = max(2, int(round(2.0 * arcLength)))
    middle = steps // 2
    samples = np.arange(0.0, arc[-1], 0.5 * spacing)
    x, z = np.interp(samples, arc, front[:, 0]), np.interp(samples, arc, front[:, 1])
    slope = (x[steps:] - x[:-steps]) / (samples[steps:] - samples[:-steps])
    raised = np.flatnonzero(z[middle:middle + len(slope)] > stillWaterLevel + 0.5 * spacing)
    if raised.size == 0:
        return ContourFace(crest, height, missing, missing, window)
    steepest = int(raised[np.argmin(slope[raised])])
    return ContourFace(crest, height, float(slope[steepest]), float(x[steepest + middle]), window)


def contourFace(positions: np.ndarray, spacing: float, crestX: float, bathymetry, reach: float, windowDepth: float,
                volumes: np.ndarray | None = None, arcLength: float = 2.0) -> ContourFace:
    '''The crest and front face's least dx/ds from the smoothed surface of `positions` (waveBreaking.py:471-557).'''
    bounds = contourWindow(spacing, crestX, bathymetry, reach, windowDepth)
    if bounds is None:
        missing = float('nan')
        return ContourFace(missing, missing, missing, missing, None)
    window = particleSurface(positions, spacing, bounds[0], bounds[1], volumes = volumes)
    return faceOfSurface(window, spacing, bathymetry.stillWaterLevel, arcLength)


@dataclass
class BreakingPoint:

    '''

    Where the wave breaks (waveBreaking.py:658-707): the time, lean, crest position and height above
    still water, the still water depth under the crest, where the face is steepest ('contour') and
    the depth there, and the trough's height.

    '''

    time: float
    lean: float
    crestX: float
    crestHeight: float
    depth: float
    troughHeight: float = float('nan')
    faceX: float = float('nan')
    faceDepth: float = float('nan')

    @property
    def breakingIndex(self) -> float:
        '''Hb / hb.'''
        return self.crestHeight / self.depth if self.depth > 0.0 else float('inf')

    @property
    def waveHeight(self) -> float:
        return self.crestHeight - self.troughHeight


@dataclass
class BreakingTracker:

    '''

    The time history of a wave running up a beach: crest position and height, face lean or slope,
    waterline and runup, and the breaking point, with the crest shoreward of `searchFromX`
    (waveBreaking.py:709-906).

    'cells': the breaking point is the first of `persistence` consecutive updates with the face
    leaning past `leanThreshold` spacings. 'contour': the start of the longest run of updates with
    the face vertical (dx/ds <= 0 over `faceArc` spacings of arc, in a window reaching `windowDepth`
    below still water, a tenth of `faceReach` by default), until a pocket of at least `pocketArea`
    spacings squared closes under a vertical face over air the last surface held: that is the jet's
    touchdown, which fixes the breaking point at the start of the run in progress and keeps the last
    open surface as `touchdownSurface`.

    '''

    bathymetry: object
    spacing: float
    leanThreshold: float = 1.0
    faceReach: float = 1.0
    searchFromX: float = -np.inf
    persistence: int = 3
    faceMeasure: str = 'cells'
    windowDepth: float | None = None
    faceArc: float = 2.0
    pocketArea: float = 4.0
    times: list = field(default_factory = list)
    crestX: list = field(default_factory = list)
    crestHeight: list = field(default_factory = list)
    lean: list = field(default_factory = list)
    faceSlope: list = field(default_factory = list)
    waterline: list = field(default_factory = list)
    breakingPoint: BreakingPoint | None = None
    maxRunup: float = float('nan')
    touchdownTime: float = float('nan')
    touchdownSurface: np.ndarray | None = None
    touchdownPocket: np.ndarray | None = None
    _candidate: BreakingPoint | None = field(default = None, repr = False)
    _streak: int = field(default = 0, repr = False)
    _vertical: bool = field(default = False, repr = False)
    _runStart: BreakingPoint | None = field(default = None, repr = False)
    _runLength: int = field(default = 0, repr = False)
    _longestRun: int = field(default = 0, repr = False)
    _lastSurface: np.ndarray | None = field(default = None, repr = False)
    _lastTime: float = field(default = float('nan'), repr = False)

    def __post_init__(self) -> None:
        if self.faceMeasure not in ('cells', 'contour'):
            raise ValueError(f"The face measure is 'cells' or 'contour', not {self.faceMeasure!r}.")

    @property
    def breakingFinal(self) -> bool:
        '''Whether the breaking point can no longer change: once recorded ('cells'); once the touchdown fixes it ('contour').'''
        if self.faceMeasure == 'contour':
            return self.touchdownSurface is not None
        return self.breakingPoint is not None

    @property
    def contourDepth(self) -> float:
        '''How far below still water the contour window reaches in deep water.'''
        return self.windowDepth if self.windowDepth is not None else 0.1 * self.faceReach

    def _point(self, time: float, lean: float, crestX: float, crestHeight: float, faceX: float,
               troughHeight: float) -> BreakingPoint:
        bathymetry = self.bathymetry
        faceDepth = float(bathymetry.depth(faceX)) if np.isfinite(faceX) else float('nan')
        return BreakingPoint(time = time, lean = lean, crestX = crestX, crestHeight = crestHeight,
                             depth = float(bathymetry.depth(crestX)), troughHeight = troughHeight, faceX = faceX,
                             faceDepth = faceDepth)

    def update(self, time: float, positions: np.ndarray, volumes: np.ndarray | None = None) -> None:
        '''Read the fluid `positions` (N, 2), with each particle's area `volumes` or s^2, at `time` (waveBreaking.py:810-906).'''
        self.observeMeasures(time, self.measures(positions, volumes))

    def crestOf(self, centers: np.ndarray, elevation: np.ndarray) -> float:
        '''The crest's x: the highest column of the free-surface profile.'''
        return float(centers[int(np.nanargmax(elevation))])

    def needsWindow(self, crestX: float) -> bool:
        '''Whether the contour measure reads its window at a profile crest at `crestX`: not while breaking cannot register.'''
        return self.faceMeasure == 'contour' and crestX >= self.searchFromX

    def measures(self, positions: np.ndarray, volumes: np.ndarray | None = None) -> dict:

        '''

        An update's measures of `positions` (N, 2): the profile's column `centers` and `elevation`,
        the waterline `tip`, and for 'cells' the face window's `occupied` cells and `windowStart`,
        for 'contour' the window's fluid fraction `grid` (x, y, phi), None without a window.

        '''

        bathymetry, s = self.bathymetry, self.spacing
        centers, elevation = freeSurfaceProfile(positions, bathymetry.startX, bathymetry.endX, s)
        crestX = self.crestOf(centers, elevation)
        measures = {'centers': centers, 'elevation': elevation, 'tip': waterlineX(positions, bathymetry, 2.0 * s)}
        if self.faceMeasure == 'cells':
            measures['occupied'], measures['windowStart'] = faceCells(positions, crestX, bathymetry.stillWaterLevel, s,
                                                                      self.faceReach)
        elif self.needsWindow(crestX):
            bounds = contourWindow(s, crestX, bathymetry, self.faceReach, self.contourDepth)
            measures['grid'] = None if bounds is None else fluidFraction(positions, s, *bounds, volumes = volumes)
        return measures

    def observeMeasures(self, time: float, measures: dict) -> None:
        '''Record one update from its measures (`measures`) and advance the breaking rule (waveBreaking.py:836-906).'''
        bathymetry, s = self.bathymetry, self.spacing
        swl = bathymetry.stillWaterLevel
        centers, elevation = measures['centers'], measures['elevation']
        crestIndex = int(np.nanargmax(elevation))
        crestX, crestHeight = float(centers[crestIndex]), float(elevation[crestIndex] - swl)
        window, faceX = None, float('nan')
        if self.faceMeasure == 'contour' and not self.needsWindow(crestX):
            lean, slope, steep = float('nan'), float('nan'), False
        elif self.faceMeasure == 'contour':
            grid = measures['grid']
            if grid is None:
                missing = float('nan')
                face = ContourFace(missing, missing, missing, missing, None)
            else:
                face = faceOfSurface(sortContours(*grid), s, swl, self.faceArc)
            if np.isfinite(face.crestX):
                crestX, crestHeight = face.crestX, face.crestHeight
            slope, faceX, window = face.slope, face.faceX, face.window
            lean, steep = float('nan'), bool(np.isfinite(slope) and slope <= 0.0)
        else:
            lean = leanOfCells(measures['occupied'], measures['windowStart'], s)
            # The lean comes in whole cells; the half-cell margin keeps rounding from deciding a tie
            slope, steep = float('nan'), lean > (self.leanThreshold - 0.5) * s
        troughHeight = troughAhead(centers, elevation, crestX, 2.0 * self.faceReach) - swl
        tip = measures['tip']
        if np.isfinite(tip):
            runup = float(bathymetry.elevation(tip) - swl)
            self.maxRunup = runup if not np.isfinite(self.maxRunup) else max(self.maxRunup, runup)
        self.times.append(time)
        self.crestX.append(crestX)
        self.crestHeight.append(crestHeight)
        self.lean.append(lean)
        self.faceSlope.append(slope)
        self.waterline.append(tip)
        if self.faceMeasure == 'cells':
            self._cellsRule(time, lean, crestX, crestHeight, faceX, troughHeight, steep)
        elif self.touchdownSurface is None:
            self._contourRule(time, lean, crestX, crestHeight, faceX, troughHeight, steep, slope, window)
        if window is not None and len(window.surface):
            self._lastSurface, self._lastTime = window.surface, time

    def _cellsRule(self, time, lean, crestX, crestHeight, faceX, troughHeight, steep) -> None:
        '''The first of `persistence` consecutive steep updates (waveBreaking.py:876-886).'''
        if self.breakingPoint is not None:
            return
        if crestX >= self.searchFromX and steep:
            if self._streak == 0:
                self._candidate = self._point(time, lean, crestX, crestHeight, faceX, troughHeight)
            self._streak += 1
            if self._streak >= self.persistence:
                self.breakingPoint = self._candidate
        else:
            self._streak, self._candidate = 0, None

    def _contourRule(self, time, lean, crestX, crestHeight, faceX, troughHeight, steep, slope, window) -> None:
        '''The touchdown, and the vertical runs before it (waveBreaking.py:887-904).'''
        s = self.spacing
        if self._vertical and window is not None and self._lastSurface is not None:
            tubes = [pocket for pocket in window.pockets if ParticleSurface.area(pocket) >= self.pocketArea * s**2
                     and airShare(pocket, self._lastSurface, s) > 0.5]
            if tubes:
                self.touchdownTime = 0.5 * (self._lastTime + time)
                self.touchdownSurface = self._lastSurface
                self.touchdownPocket = max(tubes, key = ParticleSurface.area)
                self.breakingPoint = self._runStart
        if self.touchdownSurface is None and crestX >= self.searchFromX and np.isfinite(slope):
            if steep:
                if not self._vertical:
                    self._runStart, self._runLength = self._point(time, lean, crestX, crestHeight, faceX, troughHeight), 0
                self._runLength += 1
                # The longest run so far; the earlier of two equal ones
                if self._runLength > self._longestRun:
                    self._longestRun, self.breakingPoint = self._runLength, self._runStart
            self._vertical = steep


def trackInterval(depth: float, gravity: float = 9.80665, share: float = 0.01) -> float:
    '''The tracker's reading interval, `share` sqrt(d / g) of simulated time (BreakingWave.py:514, 881).'''
    return share * math.sqrt(depth / gravity)
