'''

The breaking read-out of a 2D particle cloud, after pySurf's `waveBreaking`
(src/pysurf/waveBreaking.py:129-205, 559-657, 658-906): the free-surface profile, the trough ahead
of the crest, the lean of the crest's front face read in cells one spacing square, the waterline,
and the tracker that records the breaking point where the front face first leans forward of
vertical by a spacing over `persistence` consecutive updates (Grilli et al. 1997: breaking where
the front face first turns vertical).

The reference tracker can instead read the face on a kernel-smoothed contour of the fluid fraction
(faceMeasure 'contour', waveBreaking.py:206-557); that read-out is not built here, and asking for
it raises ValueError.

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


@dataclass
class BreakingPoint:

    '''Where the wave breaks (waveBreaking.py:658-707): the time, lean, crest position and height above still water, and the still water depth under the crest.'''

    time: float
    lean: float
    crestX: float
    crestHeight: float
    depth: float
    troughHeight: float = float('nan')

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

    The time history of a wave running up a beach: crest position and height, face lean,
    waterline and runup, and the breaking point: the first of `persistence` consecutive updates
    with the crest shoreward of `searchFromX` and the face leaning past `leanThreshold` spacings
    (waveBreaking.py:709-906, faceMeasure 'cells').

    '''

    bathymetry: object
    spacing: float
    leanThreshold: float = 1.0
    faceReach: float = 1.0
    searchFromX: float = -np.inf
    persistence: int = 3
    faceMeasure: str = 'cells'
    times: list = field(default_factory = list)
    crestX: list = field(default_factory = list)
    crestHeight: list = field(default_factory = list)
    lean: list = field(default_factory = list)
    waterline: list = field(default_factory = list)
    breakingPoint: BreakingPoint | None = None
    maxRunup: float = float('nan')
    _candidate: BreakingPoint | None = field(default = None, repr = False)
    _streak: int = field(default = 0, repr = False)

    def __post_init__(self) -> None:
        if self.faceMeasure != 'cells':
            raise ValueError("Only the 'cells' face measure is built; the smoothed-contour read-out is not.")

    @property
    def breakingFinal(self) -> bool:
        '''Whether the breaking point is recorded; read in cells it cannot change afterwards (waveBreaking.py:787-798).'''
        return self.breakingPoint is not None

    def observe(self, time: float, crestX: float, crestHeight: float, lean: float, troughHeight: float,
                tip: float) -> None:
        '''Record one update's measures and advance the breaking rule (waveBreaking.py:857-891).'''
        bathymetry, s = self.bathymetry, self.spacing
        if np.isfinite(tip):
            runup = float(bathymetry.elevation(tip) - bathymetry.stillWaterLevel)
            self.maxRunup = runup if not np.isfinite(self.maxRunup) else max(self.maxRunup, runup)
        self.times.append(time)
        self.crestX.append(crestX)
        self.crestHeight.append(crestHeight)
        self.lean.append(lean)
        self.waterline.append(tip)
        # The lean comes in whole cells; the half-cell margin keeps rounding from deciding a tie
        steep = lean > (self.leanThreshold - 0.5) * s
        if self.breakingPoint is None:
            if crestX >= self.searchFromX and steep:
                if self._streak == 0:
                    self._candidate = BreakingPoint(time = time, lean = lean, crestX = crestX, crestHeight = crestHeight,
                                                    depth = float(bathymetry.depth(crestX)), troughHeight = troughHeight)
                self._streak += 1
                if self._streak >= self.persistence:
                    self.breakingPoint = self._candidate
            else:
                self._streak, self._candidate = 0, None

    def update(self, time: float, positions: np.ndarray) -> None:
        '''Read the crest, face, trough and waterline of the fluid `positions` (N, 2) at `time` (waveBreaking.py:810-906).'''
        self.observeMeasures(time, *self.measures(positions))

    def measures(self, positions: np.ndarray) -> tuple:
        '''The profile's column centres and elevations, the face window's marked cells and start, and the waterline of `positions` (N, 2).'''
        bathymetry, s = self.bathymetry, self.spacing
        centers, elevation = freeSurfaceProfile(positions, bathymetry.startX, bathymetry.endX, s)
        crestX = self.crestOf(centers, elevation)
        occupied, windowStart = faceCells(positions, crestX, bathymetry.stillWaterLevel, s, self.faceReach)
        return centers, elevation, occupied, windowStart, waterlineX(positions, bathymetry, 2.0 * s)

    def crestOf(self, centers: np.ndarray, elevation: np.ndarray) -> float:
        '''The crest's x: the highest column of the free-surface profile.'''
        return float(centers[int(np.nanargmax(elevation))])

    def observeMeasures(self, time: float, centers: np.ndarray, elevation: np.ndarray, occupied: np.ndarray,
                        windowStart: float, tip: float) -> None:
        '''Record one update from its measures: the profile, the face window's marked cells and the waterline.'''
        swl = self.bathymetry.stillWaterLevel
        crestIndex = int(np.nanargmax(elevation))
        crestX, crestHeight = float(centers[crestIndex]), float(elevation[crestIndex] - swl)
        lean = leanOfCells(occupied, windowStart, self.spacing)
        troughHeight = troughAhead(centers, elevation, crestX, 2.0 * self.faceReach) - swl
        self.observe(time, crestX, crestHeight, lean, troughHeight, tip)


def trackInterval(depth: float, gravity: float = 9.80665, share: float = 0.01) -> float:
    '''The tracker's reading interval, `share` sqrt(d / g) of simulated time (BreakingWave.py:514, 881).'''
    return share * math.sqrt(depth / gravity)
