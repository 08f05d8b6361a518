'''

The breaking read-out measured on the device. A batch's trackers take their measures from the
engine's state where it lies: the column tops of the free-surface profile (sph_surface_profile),
then, at each run's crest, the face window's marked cells and the most landward particles near the
bed (sph_face_cells). Only those small arrays cross to the host, in place of every run's positions
at every reading (the reference's tracker reads the full arrays, waveBreaking.py:810-906).

The measures are the ones BreakingTracker.measures takes from the full positions, bit for bit:

  - a column's top is the maximum of the same float64 values, and the elevation adds the same half
    column to it on the host;
  - a cell is marked by the same floor of the same float64 quotients; the occupancy keeps a fixed
    number of bands above still water, one more than the highest column top needs, and a run whose
    particles mark a band beyond it (a particle outside the profile's columns can) is measured on
    the host instead;
  - the waterline is decided among the 8 most landward particles within the wet height of the bed,
    widened by 1e-9 of a metre against the device's rounding of the bed's interpolation, by the
    host's own test on them (waterlineFromLandward); when fewer than 3 of them pass it while more
    candidates exist, the run is measured on the host.

'''

from __future__ import annotations

import numpy as np
import torch

from pysurf_amd import _native
from pysurf_amd.breaking import faceWindow, waterlineFromLandward

LANDWARD = 8


def trackedPositions(positions: np.ndarray, isFluid: np.ndarray, dimensions: int, stripCentre: float = 0.0,
                     stripHalfWidth: float = np.inf) -> np.ndarray:
    '''The (x, z) a tracker reads: the fluid particles, in 3D those with |y - stripCentre| < stripHalfWidth.'''
    fluid = positions[isFluid]
    if dimensions == 2:
        return fluid[:, :2]
    return fluid[np.abs(fluid[:, 1] - stripCentre) < stripHalfWidth][:, [0, 2]]


def occupancyRows(words: np.ndarray, offset: int, bands: int, columns: int) -> np.ndarray:
    '''A run's (bands, columns) cells out of sph_face_cells' words: bit c % 32 of word c // 32 of each band's row.'''
    perRow = -(-columns // 32)
    block = np.ascontiguousarray(words[offset:offset + bands * perRow], dtype = '<u4')
    bits = np.unpackbits(block.view(np.uint8), bitorder = 'little')
    return bits.reshape(bands, perRow * 32)[:, :columns].astype(bool)


class DeviceReadout:

    '''

    Reads trackers' measures from a BatchEngine's state. `beds` keeps each tracker bed's polyline
    on the device, uploaded once per bed.

    '''

    def __init__(self, device = None) -> None:
        self.device = device
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

        The measures of `entries`, (run, tracker, stripCentre, stripHalfWidth) with run the index in
        `batch`: {run: (centers, elevation, occupied, windowStart, tip)}, or None for a run with no
        tracked particle. `bandMargin` is the bands kept beyond what the highest column top needs.

        '''

        engine = batch.engine
        runs = len(batch.cases)
        dims = batch.dimensions
        profiles = [_native.ProfileSpec() for _ in range(runs)]
        columns: dict = {}
        offset = 0
        for run, tracker, centre, halfWidth in entries:
            bed, s = tracker.bathymetry, tracker.spacing
            count = max(1, int(np.ceil((bed.endX - bed.startX) / s)))
            profiles[run] = _native.ProfileSpec(bed.startX, s, count, offset, centre, halfWidth)
            columns[run] = (offset, count)
            offset += count
        tops = engine.surfaceProfile(profiles)
        faces = [_native.FaceSpec() for _ in range(runs)]
        measures: dict = {}
        host: list = []
        windows: dict = {}
        word = 0
        for run, tracker, centre, halfWidth in entries:
            bed, s = tracker.bathymetry, tracker.spacing
            first, count = columns[run]
            top = tops[first:first + count]
            if not np.isfinite(top).any():
                host.append(run)
                continue
            centers = bed.startX + (np.arange(count) + 0.5) * s
            elevation = np.where(np.isfinite(top), top + 0.5 * s, np.nan)
            crestX = tracker.crestOf(centers, elevation)
            windowStart, windowEnd, base, nColumns = faceWindow(crestX, bed.stillWaterLevel, s, tracker.faceReach)
            highest = float(np.max(top[np.isfinite(top)]))
            bands = max(1, int(np.floor((highest - 0.5 * s - base) / s)) + 2 + bandMargin)
            _, bedX, bedZ = self.bedOf(bed, engine.device)
            faces[run] = _native.FaceSpec(windowStart, windowEnd, base, s, nColumns, bands, word, 1, 2.0 * s,
                                          bedX.data_ptr(), bedZ.data_ptr(), bedX.numel(), centre, halfWidth)
            windows[run] = (centers, elevation, windowStart, nColumns, bands, word)
            word += bands * -(-nColumns // 32)
        if windows:
            occupancy, overflow, landward = engine.faceCells(faces)
            for run, tracker, centre, halfWidth in entries:
                if run not in windows:
                    continue
                centers, elevation, windowStart, nColumns, bands, first = windows[run]
                candidates = landward[run]
                complete = int(np.isfinite(candidates[:, 0]).sum()) < LANDWARD
                tip = waterlineFromLandward(candidates, tracker.bathymetry, 2.0 * tracker.spacing, complete)
                if overflow[run] or tip is None:
                    host.append(run)
                    continue
                occupied = occupancyRows(occupancy, first, bands, nColumns)
                measures[run] = (centers, elevation, occupied, windowStart, tip)
        if host:
            self.fallbacks += len(host)
            positions = engine.download('position', dims)
            fluid = engine.download('fluid').astype(bool)
            byRun = {entry[0]: entry for entry in entries}
            for run in host:
                _, tracker, centre, halfWidth = byRun[run]
                a, b = batch.offsets[run], batch.offsets[run + 1]
                tracked = trackedPositions(positions[a:b], fluid[a:b], dims, centre, halfWidth)
                measures[run] = tracker.measures(tracked) if len(tracked) else None
        return measures


def hostMeasures(batch, entries: list) -> dict:
    '''The same measures from the downloaded positions, for comparison: {run: measures or None}.'''
    positions = batch.engine.download('position', batch.dimensions)
    fluid = batch.engine.download('fluid').astype(bool)
    measures = {}
    for run, tracker, centre, halfWidth in entries:
        a, b = batch.offsets[run], batch.offsets[run + 1]
        tracked = trackedPositions(positions[a:b], fluid[a:b], batch.dimensions, centre, halfWidth)
        measures[run] = tracker.measures(tracked) if len(tracked) else None
    return measures

