'''

The breaking read-out of a 2D particle cloud: what pySurf's `waveBreaking` measures
(src/pysurf/waveBreaking.py:129-657) and the tracker that turns the measures into a breaking point
(waveBreaking.py:658-906; Grilli et al. 1997: a wave breaks where its front face first turns
vertical).

The read-out has three layers:

  measures   the free-surface profile by columns, the trough ahead of the crest, the waterline,
             and the crest's front face read one of two ways:
               'cells'    the face window binned into cells one spacing square; the face's lean is
                          how far a band's face stands shoreward of the face of a band below it
               'contour'  the one-half contour of the kernel-smoothed fluid fraction, sorted into
                          the surface, pockets of air and drops of water; the face's least dx/ds
                          along its arc
  readings   one update's crest, trough, lean or slope and waterline (`BreakingTracker.update`,
             or `observeMeasures` with measures taken on the device)
  rules      'cells': breaking at the first of `persistence` consecutive readings with the face
             leaning forward by a spacing; 'contour': at the start of the last run of readings with
             the face vertical before the jet's touchdown (a pocket of air closing under it), or of
             the longest such run without one

A particle stands for the cell of water around it, so elevations include the half spacing above the
topmost particle centres.

'''

from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field

import numpy as np

MISSING = float('nan')


# -- The profile, the trough and the waterline -- #

def freeSurfaceProfile(positions: np.ndarray, x0: float, x1: float, columnWidth: float) -> tuple[np.ndarray, np.ndarray]:
    '''Each column's centre and the top of its highest particle's cell (the particle plus half a column), NaN where a column holds none (waveBreaking.py:129-162).'''
    count = max(1, int(np.ceil((x1 - x0) / columnWidth)))
    centers = x0 + (np.arange(count) + 0.5) * columnWidth
    elevation = np.full(count, np.nan)
    points = np.asarray(positions, dtype = float).reshape(-1, 2)
    column = np.floor((points[:, 0] - x0) / columnWidth).astype(np.int64)
    inRange = (column >= 0) & (column < count)
    if inRange.any():
        order = np.argsort(column[inRange], kind = 'stable')
        column, height = column[inRange][order], points[inRange, 1][order]
        starts = np.flatnonzero(np.r_[True, column[1:] != column[:-1]])
        tops = np.maximum.reduceat(height, starts)
        elevation[column[starts]] = np.where(np.isfinite(tops), tops + 0.5 * columnWidth, np.nan)
    return centers, elevation


def troughAhead(centers: np.ndarray, elevation: np.ndarray, crestX: float, reach: float = np.inf) -> float:

    '''

    The lowest free surface the crest runs into (waveBreaking.py:164-204): over the wet columns
    from the crest up to the first dry one or `reach` past it, each column's elevation replaced by
    the median of it and its two neighbours (an end column's missing neighbour mirrored from the
    other side), so that one column catching part of a particle column does not decide. NaN when
    the column just ahead of the crest is dry.

    '''

    centers, levels = np.asarray(centers), np.asarray(elevation, dtype = float)
    ahead = np.flatnonzero((centers > crestX) & (centers <= crestX + reach))
    dry = np.flatnonzero(~np.isfinite(levels[ahead]))
    run = levels[ahead[:dry[0]] if dry.size else ahead]
    if run.size == 0:
        return MISSING
    if run.size < 3:
        return float(run.max())
    before, after = np.r_[run[1], run[:-1]], np.r_[run[1:], run[-2]]
    middle = np.maximum(np.minimum(before, run), np.minimum(np.maximum(before, run), after))
    return float(middle.min())


def _landward(x: np.ndarray, z: np.ndarray, bathymetry, wetHeight: float, rank: int) -> float | None:
    '''The x of the `rank`-th most landward point within `wetHeight` of the bed, None when fewer are.'''
    wet = x[z - bathymetry.elevation(x) < wetHeight]
    if wet.size < rank:
        return None
    return float(np.partition(wet, wet.size - rank)[wet.size - rank])


def waterlineX(positions: np.ndarray, bathymetry, wetHeight: float, minimumParticles: int = 3) -> float:
    '''Where the water meets the beach: the `minimumParticles`-th most landward particle within `wetHeight` of the bed, NaN when too few are (waveBreaking.py:636-655).'''
    tip = _landward(positions[:, 0], positions[:, 1], bathymetry, wetHeight, minimumParticles)
    return MISSING if tip is None else tip


def waterlineFromLandward(landward: np.ndarray, bathymetry, wetHeight: float, complete: bool,
                          minimumParticles: int = 3) -> float | None:

    '''

    waterlineX decided among a few candidates: `landward` (m, 2) holds, by descending x, the
    particles of largest x among a set containing every particle within `wetHeight` of the bed
    (rows of NaN pad it), and `complete` says whether it holds that whole set. When
    `minimumParticles` candidates pass the wet test, or the candidates are all there are, they
    decide; otherwise returns None, and the full positions must.

    '''

    candidates = landward[np.isfinite(landward[:, 0])]
    tip = _landward(candidates[:, 0], candidates[:, 1], bathymetry, wetHeight, minimumParticles)
    if tip is not None:
        return tip
    return MISSING if complete else None


# -- The face in cells ('cells') -- #

def faceWindow(crestX: float, stillWaterLevel: float, spacing: float, reach: float,
               shoreLimit: float = np.inf) -> tuple[float, float, float, int]:
    '''The face window (waveBreaking.py:575-590): its start half a spacing behind the crest, its end `reach` ahead (at most `shoreLimit`), the base of its bands half a spacing above still water, and its number of columns.'''
    start = crestX - 0.5 * spacing
    end = min(crestX + reach, shoreLimit)
    return start, end, stillWaterLevel + 0.5 * spacing, max(1, int(np.ceil((end - start) / spacing)))


def faceCells(positions: np.ndarray, crestX: float, stillWaterLevel: float, spacing: float, reach: float,
              shoreLimit: float = np.inf) -> tuple[np.ndarray, float]:

    '''

    The face window's marked cells, (bands, columns) upward from its base, and the window's start
    (waveBreaking.py:591-612). A particle's own cell, a spacing square around it, overlaps 2 x 2
    cells of the window; it marks all four.

    '''

    start, end, base, columns = faceWindow(crestX, stillWaterLevel, spacing, reach, shoreLimit)
    x, z = positions[:, 0], positions[:, 1]
    near = (x >= start - 0.5 * spacing) & (x < end + 0.5 * spacing) & (z > base - 0.5 * spacing)
    if not near.any():
        return np.zeros((0, columns), dtype = bool), start
    column = np.floor((x[near] - 0.5 * spacing - start) / spacing).astype(np.int64)
    band = np.floor((z[near] - 0.5 * spacing - base) / spacing).astype(np.int64)
    occupied = np.zeros((int(band.max()) + 2, columns), dtype = bool)
    marksColumn = (column[:, None] + np.array([0, 1, 0, 1])).ravel()
    marksBand = (band[:, None] + np.array([0, 0, 1, 1])).ravel()
    inside = (marksColumn >= 0) & (marksColumn < columns) & (marksBand >= 0)
    occupied[marksBand[inside], marksColumn[inside]] = True
    return occupied, start


def leanOfCells(occupied: np.ndarray, windowStart: float, spacing: float) -> float:

    '''

    How far the face leans forward, from the face window's cells (waveBreaking.py:613-634). A band's
    face is the shoreward end of its run of marked cells that starts in the crest's column or the
    next; a band with no such run, or whose run reaches the window's end, has none. The lean is the
    most a band's face stands shoreward of the face of any band below it, -inf with fewer than two
    faces.

    '''

    occupied = np.asarray(occupied, dtype = bool)
    if occupied.shape[0] == 0:
        return -np.inf
    first = np.where(occupied[:, 0], 0, 1)
    begun = occupied[:, :2].any(axis = 1)
    gaps = ~occupied & (np.arange(occupied.shape[1])[None, :] >= first[:, None])
    ended = gaps.any(axis = 1)
    faces = windowStart + gaps.argmax(axis = 1)[begun & ended] * spacing
    if faces.size < 2:
        return -np.inf
    return float(np.max(faces[1:] - np.minimum.accumulate(faces)[:-1]))


def frontFaceLean(positions: np.ndarray, crestX: float, stillWaterLevel: float, spacing: float, reach: float,
                  shoreLimit: float = np.inf) -> float:
    '''How far the crest's shoreward face leans forward of vertical: leanOfCells of faceCells (waveBreaking.py:559-634).'''
    return leanOfCells(*faceCells(positions, crestX, stillWaterLevel, spacing, reach, shoreLimit), spacing)


# -- The smoothed surface ('contour') -- #

def cubicSpline2D(distance: np.ndarray, h: float) -> np.ndarray:
    '''The 2D cubic spline kernel W(r, h), normalised by 10 / (7 pi h^2), its support 2h (sphKernels.py:267-303).'''
    q = np.asarray(distance, dtype = float) / h
    norm = 10.0 / (7.0 * math.pi * h * h)
    return np.piecewise(q, [q < 1.0, (q >= 1.0) & (q < 2.0)],
                        [lambda r: norm * (1.0 - 1.5 * r**2 + 0.75 * r**3), lambda r: norm * 0.25 * (2.0 - r)**3, 0.0])


def fractionGrid(spacing: float, lower, upper) -> tuple[np.ndarray, np.ndarray]:
    '''The nodes the fluid fraction is read on: x and y every half spacing from `lower` to past `upper` (waveBreaking.py:240-243).'''
    step = 0.5 * spacing
    return tuple(np.arange(lower[axis], upper[axis] + step, step) for axis in (0, 1))


def fluidFraction(positions: np.ndarray, spacing: float, lower, upper, smoothingRatio: float = 1.3,
                  volumes: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray, np.ndarray]:

    '''

    The kernel-smoothed fluid fraction phi(x) = sum_j V_j W(|x - x_j|, h) on the fractionGrid nodes
    (waveBreaking.py:206-262), h = `smoothingRatio` spacings, V_j each particle's area (`volumes`,
    else s^2). phi is one inside the water and one half where a lattice ends, so its one-half
    contour is the free surface. Returns the nodes' x, their y and phi, shaped (ny, nx).

    The particles of the window widened by 2h each sit in the bin of the node at or below them.
    The sum goes stencil offset by stencil offset (rows outer), each offset adding its
    contributions per node in particle order, so any other implementation of the same order (the
    device's) agrees with it to rounding.

    '''

    step, h = 0.5 * spacing, smoothingRatio * spacing
    gridX, gridY = fractionGrid(spacing, lower, upper)
    shape = (len(gridY), len(gridX))
    fraction = np.zeros(shape)
    points = np.asarray(positions, dtype = float).reshape(-1, 2)
    near = np.all((points > np.asarray(lower[:2]) - 2.0 * h) & (points < np.asarray(upper[:2]) + 2.0 * h), axis = 1)
    points = points[near]
    areas = np.full(len(points), spacing**2) if volumes is None else np.asarray(volumes, dtype = float)[near]
    binX = np.floor((points[:, 0] - gridX[0]) / step).astype(np.int64)
    binY = np.floor((points[:, 1] - gridY[0]) / step).astype(np.int64)
    reach = int(np.ceil(2.0 * h / step))
    for dy, dx in itertools.product(range(-reach, reach + 2), repeat = 2):
        nodeX, nodeY = binX + dx, binY + dy
        on = (nodeX >= 0) & (nodeX < shape[1]) & (nodeY >= 0) & (nodeY < shape[0])
        gap = np.hypot(gridX[nodeX[on]] - points[on, 0], gridY[nodeY[on]] - points[on, 1])
        added = np.bincount(nodeY[on] * shape[1] + nodeX[on], weights = areas[on] * cubicSpline2D(gap, h),
                            minlength = fraction.size)
        fraction += added.reshape(shape)
    return gridX, gridY, fraction


def halfContours(gridX: np.ndarray, gridY: np.ndarray, fraction: np.ndarray) -> list:
    '''The one-half contours of a gridded fraction as polylines, traced by contourpy (waveBreaking.py:300-308).'''
    import contourpy
    tracer = contourpy.contour_generator(gridX, gridY, fraction, line_type = contourpy.LineType.Separate)
    return [np.asarray(line) for line in tracer.lines(0.5)]


def fluidSurface(positions: np.ndarray, spacing: float, lower, upper, smoothingRatio: float = 1.3,
                 volumes: np.ndarray | None = None) -> list:
    '''The one-half contours of the smoothed fluid fraction over the window [lower, upper] (waveBreaking.py:264-298).'''
    return halfContours(*fluidFraction(positions, spacing, lower, upper, smoothingRatio, volumes))


@dataclass
class ParticleSurface:

    '''

    The free surface within a window (waveBreaking.py:310-340): `surface`, the longest contour that
    ends on the window's edges, running from its end of least x ((0, 2) when there is none);
    `pockets`, closed contours around air; `drops`, closed contours around water.

    '''

    surface: np.ndarray
    pockets: list
    drops: list

    @staticmethod
    def area(loop: np.ndarray) -> float:
        '''The area a closed contour encloses, by the shoelace formula.'''
        x, y = loop[:, 0], loop[:, 1]
        return 0.5 * abs(float(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1))))


def _closed(line: np.ndarray) -> bool:
    '''Whether a contour closes on itself.'''
    return len(line) >= 3 and bool(np.allclose(line[0], line[-1]))


def _enclosedFraction(loop: np.ndarray, gridX: np.ndarray, gridY: np.ndarray, fraction: np.ndarray) -> float:
    '''The mean fraction of the nodes inside a closed contour; for a loop too small to hold one, the fraction at the node nearest its vertex mean.'''
    from matplotlib.path import Path
    low, high = loop.min(axis = 0), loop.max(axis = 0)
    columns = np.flatnonzero((gridX >= low[0]) & (gridX <= high[0]))
    rows = np.flatnonzero((gridY >= low[1]) & (gridY <= high[1]))
    nodeX, nodeY = np.meshgrid(gridX[columns], gridY[rows])
    inside = Path(loop).contains_points(np.column_stack([nodeX.ravel(), nodeY.ravel()]))
    if inside.any():
        return float(np.mean(fraction[np.ix_(rows, columns)].ravel()[inside]))
    centre = loop.mean(axis = 0)
    return float(fraction[int(np.argmin(np.abs(gridY - centre[1]))), int(np.argmin(np.abs(gridX - centre[0])))])


def sortContours(gridX: np.ndarray, gridY: np.ndarray, fraction: np.ndarray) -> ParticleSurface:

    '''

    The one-half contours of a gridded fraction sorted into the surface, pockets and drops
    (waveBreaking.py:342-402). An open contour is part of the surface the window cuts, and the
    longest (the first of equal length) is kept; a closed one holds air when the mean fraction of
    the nodes inside it is below one half.

    '''

    surface, pockets, drops = np.zeros((0, 2)), [], []
    for line in halfContours(gridX, gridY, fraction):
        if not _closed(line):
            if len(line) > len(surface):
                surface = line if line[0, 0] <= line[-1, 0] else line[::-1]
        elif _enclosedFraction(line, gridX, gridY, fraction) < 0.5:
            pockets.append(line)
        else:
            drops.append(line)
    return ParticleSurface(surface = surface, pockets = pockets, drops = drops)


def particleSurface(positions: np.ndarray, spacing: float, lower, upper, smoothingRatio: float = 1.3,
                    volumes: np.ndarray | None = None) -> ParticleSurface:
    '''The surface, pockets and drops of the particles' smoothed fraction over the window [lower, upper] (waveBreaking.py:342-402).'''
    return sortContours(*fluidFraction(positions, spacing, lower, upper, smoothingRatio, volumes))


def airShare(loop: np.ndarray, surface: np.ndarray, spacing: float) -> float:

    '''

    The share of a closed contour's area on the air side of an open surface running from its end
    of least x (waveBreaking.py:404-443): outside the polygon the surface makes when closed through
    the water a spacing below everything. The loop's inside is sampled on nodes a quarter spacing
    apart, or at its vertex mean when it holds none. NaN for a surface of fewer than two points.

    '''

    from matplotlib.path import Path
    if len(surface) < 2:
        return MISSING
    below = min(float(surface[:, 1].min()), float(loop[:, 1].min())) - spacing
    water = Path(np.vstack([surface, [[surface[-1, 0], below], [surface[0, 0], below]]]))
    step = 0.25 * spacing
    low, high = loop.min(axis = 0), loop.max(axis = 0)
    nodeX, nodeY = np.meshgrid(np.arange(low[0], high[0] + step, step), np.arange(low[1], high[1] + step, step))
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

    @classmethod
    def unread(cls, window: ParticleSurface | None = None, crestX: float = MISSING,
               crestHeight: float = MISSING) -> ContourFace:
        '''A face with no slope read.'''
        return cls(crestX, crestHeight, MISSING, MISSING, window)


def contourWindow(spacing: float, crestX: float, bathymetry, reach: float, windowDepth: float):

    '''

    The window the face is read in (waveBreaking.py:520-530), (lower, upper) corners, or None when
    less than a spacing of depth or two spacings of length remain. Along x from `reach` behind the
    crest to twice `reach` ahead, stopping where the bed first comes within three spacings of the
    lower edge (the water's edge against the bed must not read as surface); up from `windowDepth`
    below still water (in shallow water, half of the depth under the crest left after three
    spacings) to twice `reach` above it.

    '''

    level = bathymetry.stillWaterLevel
    depth = min(windowDepth, 0.5 * (float(bathymetry.depth(crestX)) - 3.0 * spacing))
    if depth < spacing:
        return None
    bottom = level - depth
    stations = np.linspace(crestX, min(crestX + 2.0 * reach, bathymetry.endX), 400)
    grounded = np.flatnonzero(bathymetry.elevation(stations) > bottom - 3.0 * spacing)
    end = float(stations[grounded[0]] if grounded.size else stations[-1])
    if end <= crestX + 2.0 * spacing:
        return None
    return np.array([crestX - reach, bottom]), np.array([end, level + 2.0 * reach])


def faceOfSurface(window: ParticleSurface, spacing: float, stillWaterLevel: float, arcLength: float = 2.0) -> ContourFace:

    '''

    The crest, the surface's highest point, and the least dx/ds along the front face from it to
    the surface's shoreward end (waveBreaking.py:531-557). The face is resampled every half spacing
    of arc; its slope at a sample is the change in x over the `arcLength` spacings of arc centred
    there, read only where the face stands more than half a spacing above still water.

    '''

    surface = window.surface
    if len(surface) < 8:
        return ContourFace.unread(window)
    top = int(np.argmax(surface[:, 1]))
    crest, height = float(surface[top, 0]), float(surface[top, 1] - stillWaterLevel)
    front = surface[top:]
    arc = np.r_[0.0, np.cumsum(np.hypot(*np.diff(front, axis = 0).T))]
    if arc[-1] < max(4.0, arcLength + 2.0) * spacing:
        return ContourFace.unread(window, crest, height)
    span = max(2, int(round(2.0 * arcLength)))
    centre = span // 2
    stations = np.arange(0.0, arc[-1], 0.5 * spacing)
    x, z = np.interp(stations, arc, front[:, 0]), np.interp(stations, arc, front[:, 1])
    dxds = (x[span:] - x[:-span]) / (stations[span:] - stations[:-span])
    raised = np.flatnonzero(z[centre:centre + dxds.size] > stillWaterLevel + 0.5 * spacing)
    if raised.size == 0:
        return ContourFace.unread(window, crest, height)
    steepest = int(raised[np.argmin(dxds[raised])])
    return ContourFace(crest, height, float(dxds[steepest]), float(x[steepest + centre]), window)


def contourFace(positions: np.ndarray, spacing: float, crestX: float, bathymetry, reach: float, windowDepth: float,
                volumes: np.ndarray | None = None, arcLength: float = 2.0) -> ContourFace:
    '''The crest and the front face's least dx/ds from the smoothed surface of `positions` around `crestX` (waveBreaking.py:471-557).'''
    bounds = contourWindow(spacing, crestX, bathymetry, reach, windowDepth)
    if bounds is None:
        return ContourFace.unread()
    return faceOfSurface(particleSurface(positions, spacing, *bounds, volumes = volumes), spacing,
                         bathymetry.stillWaterLevel, arcLength)


# -- The tracker -- #

@dataclass
class BreakingPoint:

    '''

    Where the wave breaks (waveBreaking.py:658-707): the time, lean, crest position and height above
    still water, the still water depth under the crest, the trough's height, where the face is
    steepest ('contour') and the depth there.

    '''

    time: float
    lean: float
    crestX: float
    crestHeight: float
    depth: float
    troughHeight: float = MISSING
    faceX: float = MISSING
    faceDepth: float = MISSING

    @property
    def breakingIndex(self) -> float:
        '''Hb / hb.'''
        return self.crestHeight / self.depth if self.depth > 0.0 else float('inf')

    @property
    def waveHeight(self) -> float:
        return self.crestHeight - self.troughHeight


@dataclass
class _Reading:
    '''One update's read-out, as the rules see it.'''

    time: float
    crestX: float
    crestHeight: float
    troughHeight: float
    lean: float = MISSING
    slope: float = MISSING
    faceX: float = MISSING
    window: ParticleSurface | None = None
    steep: bool = False
    searched: bool = False


class _FirstSteepRun:

    '''The 'cells' rule (waveBreaking.py:876-886): the first reading of the first `persistence` consecutive steep ones.'''

    def __init__(self, persistence: int) -> None:
        self.persistence = persistence
        self.streak = 0
        self.candidate: BreakingPoint | None = None
        self.lastSurface = None

    def observe(self, tracker: BreakingTracker, reading: _Reading) -> None:
        if tracker.breakingPoint is not None:
            return
        if not (reading.searched and reading.steep):
            self.streak, self.candidate = 0, None
            return
        if self.streak == 0:
            self.candidate = tracker._pointOf(reading)
        self.streak += 1
        if self.streak >= self.persistence:
            tracker.breakingPoint = self.candidate


class _VerticalRunsToTouchdown:

    '''

    The 'contour' rule (waveBreaking.py:887-904). A run is a stretch of consecutive readings with
    the face vertical; the breaking point is the start of the longest run so far (the earlier of
    equal ones) until the touchdown, a pocket of at least `pocketArea` spacings squared lying
    mostly on the air side of the last surface read, found right after a vertical reading; the
    touchdown fixes the breaking point at the start of the run in progress.

    '''

    def __init__(self, pocketArea: float) -> None:
        self.pocketArea = pocketArea
        self.vertical = False
        self.runStart: BreakingPoint | None = None
        self.runLength = 0
        self.longest = 0
        self.lastSurface: np.ndarray | None = None
        self.lastTime = MISSING

    def _touchdown(self, tracker: BreakingTracker, reading: _Reading) -> None:
        if not (self.vertical and reading.window is not None and self.lastSurface is not None):
            return
        s = tracker.spacing
        tubes = [pocket for pocket in reading.window.pockets if ParticleSurface.area(pocket) >= self.pocketArea * s**2
                 and airShare(pocket, self.lastSurface, s) > 0.5]
        if tubes:
            tracker.touchdownTime = 0.5 * (self.lastTime + reading.time)
            tracker.touchdownSurface = self.lastSurface
            tracker.touchdownPocket = max(tubes, key = ParticleSurface.area)
            tracker.breakingPoint = self.runStart

    def observe(self, tracker: BreakingTracker, reading: _Reading) -> None:
        if tracker.touchdownSurface is None:
            self._touchdown(tracker, reading)
            if tracker.touchdownSurface is None and reading.searched and np.isfinite(reading.slope):
                if reading.steep:
                    if not self.vertical:
                        self.runStart, self.runLength = tracker._pointOf(reading), 0
                    self.runLength += 1
                    if self.runLength > self.longest:
                        self.longest, tracker.breakingPoint = self.runLength, self.runStart
                self.vertical = reading.steep
        if reading.window is not None and len(reading.window.surface):
            self.lastSurface, self.lastTime = reading.window.surface, reading.time


@dataclass
class BreakingTracker:

    '''

    The time history of a wave running up a beach, read every update: crest position and height,
    face lean ('cells') or least slope ('contour'), waterline and highest runup, and the breaking
    point, counted only with the crest shoreward of `searchFromX` (waveBreaking.py:709-906).

    'cells': breaking at the first of `persistence` consecutive readings with the face leaning
    past `leanThreshold` spacings. 'contour': the face is vertical where dx/ds <= 0 over `faceArc`
    spacings of arc, in a window reaching `windowDepth` below still water (a tenth of `faceReach`
    by default); breaking at the start of the longest run of vertical readings, fixed at the start
    of the run in progress when a pocket of at least `pocketArea` spacings squared closes under a
    vertical face over air the last surface held: the jet's touchdown, whose time and last open
    surface the tracker keeps.

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
    maxRunup: float = MISSING
    touchdownTime: float = MISSING
    touchdownSurface: np.ndarray | None = None
    touchdownPocket: np.ndarray | None = None
    _rule: object = field(default = None, repr = False)

    def __post_init__(self) -> None:
        if self.faceMeasure == 'cells':
            self._rule = _FirstSteepRun(self.persistence)
        elif self.faceMeasure == 'contour':
            self._rule = _VerticalRunsToTouchdown(self.pocketArea)
        else:
            raise ValueError(f"The face measure is 'cells' or 'contour', not {self.faceMeasure!r}.")

    @property
    def breakingFinal(self) -> bool:
        '''Whether the breaking point can no longer move: once recorded ('cells'), once the touchdown fixes it ('contour').'''
        if self.faceMeasure == 'contour':
            return self.touchdownSurface is not None
        return self.breakingPoint is not None

    @property
    def lastSurface(self) -> np.ndarray | None:
        '''The last open surface the 'contour' measure read (None before one, and for 'cells').'''
        return self._rule.lastSurface

    @property
    def contourDepth(self) -> float:
        '''How far below still water the contour window reaches in deep water.'''
        return 0.1 * self.faceReach if self.windowDepth is None else self.windowDepth

    def crestOf(self, centers: np.ndarray, elevation: np.ndarray) -> float:
        '''The crest's x: the centre of the profile's highest column.'''
        return float(centers[int(np.nanargmax(elevation))])

    def needsWindow(self, crestX: float) -> bool:
        '''Whether the 'contour' measure reads its window with the profile's crest at `crestX` (not where breaking cannot count).'''
        return self.faceMeasure == 'contour' and crestX >= self.searchFromX

    def _pointOf(self, reading: _Reading) -> BreakingPoint:
        bed = self.bathymetry
        return BreakingPoint(time = reading.time, lean = reading.lean, crestX = reading.crestX,
                             crestHeight = reading.crestHeight, depth = float(bed.depth(reading.crestX)),
                             troughHeight = reading.troughHeight, faceX = reading.faceX,
                             faceDepth = float(bed.depth(reading.faceX)) if np.isfinite(reading.faceX) else MISSING)

    def update(self, time: float, positions: np.ndarray, volumes: np.ndarray | None = None) -> None:
        '''Read the fluid's (x, z) `positions`, each particle's area `volumes` (else s^2), at `time` (waveBreaking.py:810-906).'''
        self.observeMeasures(time, self.measures(positions, volumes))

    def measures(self, positions: np.ndarray, volumes: np.ndarray | None = None) -> dict:

        '''

        One update's measures of `positions` (N, 2): the profile's column `centers` and `elevation`,
        the waterline `tip`, and for 'cells' the face window's `occupied` cells and `windowStart`,
        for 'contour' (with the crest where it counts) the window's fluid fraction `grid`
        (x, y, phi), None without a window.

        '''

        bed, s = self.bathymetry, self.spacing
        centers, elevation = freeSurfaceProfile(positions, bed.startX, bed.endX, s)
        crestX = self.crestOf(centers, elevation)
        found = {'centers': centers, 'elevation': elevation, 'tip': waterlineX(positions, bed, 2.0 * s)}
        if self.faceMeasure == 'cells':
            found['occupied'], found['windowStart'] = faceCells(positions, crestX, bed.stillWaterLevel, s, self.faceReach)
        elif self.needsWindow(crestX):
            bounds = contourWindow(s, crestX, bed, self.faceReach, self.contourDepth)
            found['grid'] = None if bounds is None else fluidFraction(positions, s, *bounds, volumes = volumes)
        return found

    def _read(self, time: float, found: dict) -> _Reading:
        '''The reading the measures give (waveBreaking.py:836-875).'''
        bed, s = self.bathymetry, self.spacing
        level = bed.stillWaterLevel
        centers, elevation = found['centers'], found['elevation']
        top = int(np.nanargmax(elevation))
        crestX, crestHeight = float(centers[top]), float(elevation[top] - level)
        reading = {'searched': crestX >= self.searchFromX}
        if self.faceMeasure == 'cells':
            lean = leanOfCells(found['occupied'], found['windowStart'], s)
            # the lean comes in whole cells; the half-cell margin keeps rounding from deciding a tie
            reading.update(lean = lean, steep = lean > (self.leanThreshold - 0.5) * s)
        elif self.needsWindow(crestX):
            grid = found['grid']
            face = ContourFace.unread() if grid is None else faceOfSurface(sortContours(*grid), s, level, self.faceArc)
            if np.isfinite(face.crestX):
                crestX, crestHeight = face.crestX, face.crestHeight
            reading.update(slope = face.slope, faceX = face.faceX, window = face.window,
                           steep = bool(np.isfinite(face.slope) and face.slope <= 0.0),
                           searched = crestX >= self.searchFromX)
        trough = troughAhead(centers, elevation, crestX, 2.0 * self.faceReach) - level
        return _Reading(time = time, crestX = crestX, crestHeight = crestHeight, troughHeight = trough, **reading)

    def observeMeasures(self, time: float, measures: dict) -> None:
        '''Record one update from its measures and advance the breaking rule.'''
        reading = self._read(time, measures)
        tip = measures['tip']
        if np.isfinite(tip):
            runup = float(self.bathymetry.elevation(tip) - self.bathymetry.stillWaterLevel)
            self.maxRunup = runup if not np.isfinite(self.maxRunup) else max(self.maxRunup, runup)
        for history, value in ((self.times, time), (self.crestX, reading.crestX), (self.crestHeight, reading.crestHeight),
                               (self.lean, reading.lean), (self.faceSlope, reading.slope), (self.waterline, tip)):
            history.append(value)
        self._rule.observe(self, reading)


def trackInterval(depth: float, gravity: float = 9.80665, share: float = 0.01) -> float:
    '''The tracker's reading interval, `share` sqrt(d / g) of simulated time (BreakingWave.py:514, 881).'''
    return share * math.sqrt(depth / gravity)
