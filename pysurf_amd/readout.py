'''

The breaking read-out measured on the device. A batch's trackers take their measures from the
engine's state where it lies: the column tops of the free-surface profile (sph_surface_profile),
then, at each run's crest, the face window's marked cells and the most landward particles near the
bed (sph_face_cells) and, for a tracker reading the smoothed surface, the fluid fraction over its
window (sph_fluid_fraction). Only those small arrays cross to the host, in place of every run's
positions at every reading (the reference's tracker reads the full arrays, waveBreaking.py:810-906).

The measures are the ones BreakingTracker.measures takes from the full positions:

  - a column's top is the maximum of the same float64 values, and the elevation adds the same half
    column to it on the host: bit for bit;
  - a cell is marked by the same floor of the same float64 quotients, bit for bit; the occupancy
    keeps a fixed number of bands above still water, one more than the highest column top needs,
    and a run whose particles mark a band beyond it (a particle outside the profile's columns can)
    is measured on the host instead;
  - the waterline is decided among the 8 most landward particles within the wet height of the bed,
    widened by 1e-9 of a metre against the device's rounding of the bed's interpolation, by the
    host's own test on them (waterlineFromLandward), bit for bit; when fewer than 3 of them pass it
    while more candidates exist, the run is measured on the host;
  - the fluid fraction is the same sum in the same order of bins, each bin's particles in slot
    order, in float64; the kernel's root and powers may round differently from numpy's, so it agrees
    to rounding (1e-12), and it does not depend on the run's neighbours in the batch.

'''

from __future__ import annotations

import math

import numpy as np
import torch

from pysurf_amd import _native
from pysurf_amd.breaking import contourWindow, faceWindow, fractionGrid, waterlineFromLandward

LANDWARD = 8
SMOOTHING_RATIO = 1.3


def trackedPositions(positions: np.ndarray, isFluid: np.ndarray, dimensions: int, stripCentre: float = 0.0,
                     stripHalfWidth: float = np.inf) -> np.ndarray:
    '''The (x, z) a tracker reads: the fluid particles, in 3D those with |y - stripCentre| < stripHalfWidth.'''
    return positions[trackedMask(positions, isFluid, dimensions, stripCentre, stripHalfWidth)][:, [0, dimensions - 1]]


def trackedMask(positions: np.ndarray, isFluid: np.ndarray, dimensions: int, stripCentre: float = 0.0,
                stripHalfWidth: float = np.inf) -> np.ndarray:
    '''Which particles a tracker reads.'''
    mask = np.asarray(isFluid, dtype = bool).copy()
    if dimensions == 3:
        mask &= np.abs(positions[:, 1] - stripCentre) < stripHalfWidth
    return mask


def occupancyRows(words: np.ndarray, offset: int, bands: int, columns: int) -> np.ndarray:
    '''A run's (bands, columns) cells out of sph_face_cells' words: bit c % 32 of word c // 32 of each band's row.'''
    perRow = -(-columns // 32)
    block = np.ascontiguousarray(words[offset:offset + bands * perRow], dtype = '<u4')
    bits = np.unpackbits(block.view(np.uint8), bitorder = 'little')
    return bits.reshape(bands, perRow * 32)[:, :columns].astype(bool)


def fractionSpec(spacing: float, bounds, offset: int, areaPerMass: float, stripCentre: float,
                 stripHalfWidth: float) -> tuple:
    '''The sph_fraction_spec of a window (`bounds` = lower, upper) and its nodes' x and y.'''
    lower, upper = bounds
    gridX, gridY = fractionGrid(spacing, lower, upper)
    step = 0.5 * spacing
    h = SMOOTHING_RATIO * spacing
    spec = _native.FractionSpec()
    spec.x0, spec.y0 = gridX[0], gridY[0]
    # The nodes stand where numpy's arange puts them: the first plus i times the first difference
    spec.stepX = gridX[1] - gridX[0] if len(gridX) > 1 else step
    spec.stepY = gridY[1] - gridY[0] if len(gridY) > 1 else step
    spec.binWidth = step
    spec.nx, spec.ny = len(gridX), len(gridY)
    spec.reach = int(math.ceil(2.0 * h / step))
    spec.offset = offset
    spec.h = h
    spec.area = spacing**2
    spec.areaPerMass = areaPerMass
    spec.lowX, spec.lowZ = lower[0] - 2.0 * h, lower[1] - 2.0 * h
    spec.highX, spec.highZ = upper[0] + 2.0 * h, upper[1] + 2.0 * h
    spec.stripCentre, spec.stripHalfWidth = stripCentre, stripHalfWidth
    return spec, gridX, gridY


class DeviceReadout:

    '''

    Reads trackers' measures from a BatchEngine's state. `beds` keeps the polyline of each bed the
    last reading's trackers read on the device, uploaded once while it stays in use; `fallbacks`
    counts the runs measured on the host.

    An entry is (run, tracker, stripCentre, stripHalfWidth, areaPerMass): the run's index in the
    batch, its BreakingTracker, the strip a 3D run's tracker reads, and a particle's area in the
    smoothed surface per unit mass (0: the spacing squared).

    '''

    def __init__(self) -> None:
        self.beds: dict = {}
        self.fallbacks = 0

    def bedOf(self, bathymetry, device) -> tuple:
        key = id(bathymetry)
        if key not in self.beds:
            x = torch.as_tensor(np.ascontiguousarray(bathymetry.x, dtype = np.float64), device = device)
            z = torch.as_tensor(np.ascontiguousarray(bathymetry.z, dtype = np.float64), device = device)
            self.beds[key] = (bathymetry, x, z)
        return self.beds[key]

    def measure(self, batch, entries: list, bandMargin: int = 1) -> dict:

        '''

        The measures of `entries`: {run: the dict BreakingTracker.measures returns}, or None for a
        run with no tracked particle. `bandMargin` is the bands kept beyond what the highest column
        top needs.

        '''

        engine = batch.engine
        runs = len(batch.cases)
        profiles = [_native.ProfileSpec() for _ in range(runs)]
        columns: dict = {}
        offset = 0
        for run, tracker, centre, halfWidth, _ in entries:
            bed, s = tracker.bathymetry, tracker.spacing
            count = max(1, int(np.ceil((bed.endX - bed.startX) / s)))
            profiles[run] = _native.ProfileSpec(bed.startX, s, count, offset, centre, halfWidth)
            columns[run] = (offset, count)
            offset += count
        tops = engine.surfaceProfile(profiles)
        faces = [_native.FaceSpec() for _ in range(runs)]
        fractions = [_native.FractionSpec() for _ in range(runs)]
        measures: dict = {}
        host: list = []
        pending: dict = {}
        word = node = 0
        for run, tracker, centre, halfWidth, areaPerMass in entries:
            bed, s = tracker.bathymetry, tracker.spacing
            first, count = columns[run]
            top = tops[first:first + count]
            if not np.isfinite(top).any():
                host.append(run)
                continue
            centers = bed.startX + (np.arange(count) + 0.5) * s
            elevation = np.where(np.isfinite(top), top + 0.5 * s, np.nan)
            crestX = tracker.crestOf(centers, elevation)
            _, bedX, bedZ = self.bedOf(bed, engine.device)
            found = {'centers': centers, 'elevation': elevation}
            if tracker.faceMeasure == 'cells':
                windowStart, windowEnd, base, nColumns = faceWindow(crestX, bed.stillWaterLevel, s, tracker.faceReach)
                highest = float(np.max(top[np.isfinite(top)]))
                bands = max(1, int(np.floor((highest - 0.5 * s - base) / s)) + 2 + bandMargin)
                faces[run] = _native.FaceSpec(windowStart, windowEnd, base, s, nColumns, bands, word, 1, 2.0 * s,
                                              bedX.data_ptr(), bedZ.data_ptr(), bedX.numel(), centre, halfWidth)
                found['windowStart'] = windowStart
                pending[run] = (found, (word, bands, nColumns), None)
                word += bands * -(-nColumns // 32)
            else:
                # The landward points only
                faces[run] = _native.FaceSpec(0.0, 0.0, 0.0, s, 0, 0, 0, 1, 2.0 * s, bedX.data_ptr(), bedZ.data_ptr(),
                                              bedX.numel(), centre, halfWidth)
                grid = None
                if tracker.needsWindow(crestX):
                    bounds = contourWindow(s, crestX, bed, tracker.faceReach, tracker.contourDepth)
                    found['grid'] = None
                    if bounds is not None:
                        spec, gridX, gridY = fractionSpec(s, bounds, node, areaPerMass, centre, halfWidth)
                        fractions[run] = spec
                        grid = (gridX, gridY, node)
                        node += spec.nx * spec.ny
                pending[run] = (found, None, grid)
        if pending:
            occupancy, overflow, landward = engine.faceCells(faces)
            fraction = engine.fluidFraction(fractions) if node else None
            for run, tracker, *_ in entries:
                if run not in pending:
                    continue
                found, cells, grid = pending[run]
                candidates = landward[run]
                complete = int(np.isfinite(candidates[:, 0]).sum()) < LANDWARD
                tip = waterlineFromLandward(candidates, tracker.bathymetry, 2.0 * tracker.spacing, complete)
                if overflow[run] or tip is None:
                    host.append(run)
                    continue
                found['tip'] = tip
                if cells is not None:
                    found['occupied'] = occupancyRows(occupancy, *cells)
                if grid is not None:
                    gridX, gridY, first = grid
                    found['grid'] = (gridX, gridY, fraction[first:first + len(gridX) * len(gridY)].reshape(len(gridY), len(gridX)))
                measures[run] = found
        if host:
            self.fallbacks += len(host)
            measures.update(hostMeasures(batch, [entry for entry in entries if entry[0] in host]))
        # Only the beds of the trackers still read stay on the device
        reading = {id(entry[1].bathymetry) for entry in entries}
        self.beds = {key: bed for key, bed in self.beds.items() if key in reading}
        return measures


def hostMeasures(batch, entries: list) -> dict:
    '''The same measures from the downloaded state, measured on the host: {run: measures or None}.'''
    engine, dims = batch.engine, batch.dimensions
    positions = engine.download('position', dims)
    fluid = engine.download('fluid').astype(bool)
    masses = engine.download('mass')
    measures = {}
    for run, tracker, centre, halfWidth, areaPerMass in entries:
        a, b = batch.offsets[run], batch.offsets[run + 1]
        mask = trackedMask(positions[a:b], fluid[a:b], dims, centre, halfWidth)
        tracked = positions[a:b][mask][:, [0, dims - 1]]
        volumes = masses[a:b][mask] * areaPerMass if areaPerMass > 0.0 else None
        measures[run] = tracker.measures(tracked, volumes) if len(tracked) else None
    return measures
