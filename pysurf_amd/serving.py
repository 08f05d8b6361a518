'''

Runs stepped to their breaking point, alone or many at once, and the request loop that serves
them: Poisson arrivals, admission under an HBM budget, continuous batching.

A fleet holds the runs of one dimension in one BatchEngine. It steps them one tracking interval
at a time (0.01 sqrt(d / g) of each run's simulated time, BreakingWave.py:514, 881): each run's
device clock stops at its next reading, the trackers' measures are taken on the device for all runs
at once (pysurf_amd.readout), and each run's breaking tracker reads its own (a 3D run's tracker
reads the smoothed surface of the strip a quarter depth wide along the tank's centre line, in x-z,
over the bed's section there). A run is done when its breaking point is recorded or its time runs
out.

A run whose state stops being finite or whose time step collapses stops on the device at the step
that finds it (sph_run_failures): the request loop ends that request, recording why, and the other
runs step on; runToBreaking raises NumericalInstabilityError, as the reference's run loop does.

Membership changes between intervals: finished runs leave and admitted runs join, and the fleet
rebuilds its engine from the runs' current states, each run's particles in the same slot order and
its clock carried over; the runs that stay are copied from the old engine on the device. Because a run's step depends only on its own state and clock (its cells,
its time step and the fixed launch shape), a run's history is the same bit for bit alone and in any
mix.

'''

from __future__ import annotations

import math
import time as timeModule
from dataclasses import dataclass, field

import numpy as np
import torch

from pysurf_amd import _native, readout
from pysurf_amd.batch import BatchEngine, stateFields
from pysurf_amd.breaking import BreakingTracker, trackInterval
from pysurf_amd.errors import NumericalInstabilityError

# HBM per particle the admission counts: state 81 B (f64 position, velocity, density; f32 pressure,
# acceleration, mass, length; a flag), the sorted record and its key, G and L, the neighbor list,
# the run id, cell key, order, scattered order and (3D) eighth along x, and the resident accelerations
def bytesPerParticle(dimensions: int, listCapacity: int) -> int:
    d = dimensions
    state = 8 * d * 2 + 8 + 4 + 4 * d + 4 + 4 + 1
    return state + 64 + 4 + 8 * d * d + 4 * (listCapacity + 1) + 16 + 1 + 4 * d


def trackerOf(case) -> BreakingTracker:

    '''

    The run's breaking tracker: face reach one depth, crests offshore of the bed's toe ignored; a 2D
    run's reads the face in cells, or the smoothed surface from the exact start
    (BreakingWave.py:763-765), a 3D run's the smoothed surface of its strip over the bed's section
    there (BreakingWave3D.py:693-694).

    '''

    depth = case.parameters.get('depth', 1.0)
    toe = case.parameters.get('toeX', 0.0) * (depth if case.dimensions == 3 else 1.0)
    if case.dimensions == 2:
        # The exact start reads the smoothed surface (BreakingWave.py:765)
        return BreakingTracker(case.boundary.bathymetry, case.spacing, faceReach = depth, searchFromX = toe,
                               faceMeasure = case.parameters.get('faceMeasure', 'cells'),
                               faceArc = case.parameters.get('faceArc', 2.0))
    startY, endY = case.boundary.field.bounds[2:]
    bed = case.boundary.field.section(0.5 * (startY + endY))
    return BreakingTracker(bed, case.spacing, faceReach = depth, searchFromX = toe, faceMeasure = 'contour')


def stripOf(case) -> tuple[float, float]:
    '''The centre line and half width of the strip a 3D run's tracker reads: the tank's centre, a quarter depth wide.'''
    if case.dimensions == 2:
        return 0.0, float('inf')
    startY, endY = case.boundary.field.bounds[2:]
    return 0.5 * (startY + endY), 0.125 * case.parameters.get('depth', 1.0)


def areaPerMass(case) -> float:

    '''

    A particle's area in the smoothed surface per unit mass: in 3D its volume over the strip's width
    (BreakingWave3D.py:476-480); in 2D its volume where particles have split, and 0 (the spacing
    squared) where they have not (BreakingWave.py:908).

    '''

    if case.dimensions == 3:
        return 1.0 / (case.referenceDensity * 2.0 * stripOf(case)[1])
    return 1.0 / case.referenceDensity if case.particles.lengths is not None else 0.0


def readingEntry(index: int, member) -> tuple:
    '''What the read-out needs of a member (pysurf_amd.readout): its run, tracker, strip and area rule.'''
    return (index, member.tracker, *stripOf(member.case), areaPerMass(member.case))


@dataclass
class Member:

    '''One run in a fleet: its case (its start), its clock, its tracker and its deadlines; its state lives in the fleet's engine.'''

    key: object
    case: object
    endTime: float
    time: float = 0.0
    steps: int = 0
    tracker: BreakingTracker = None
    nextTrack: float = 0.0
    interval: float = 0.0
    done: bool = False
    finishedWall: float = float('nan')
    doneTime: float = float('nan')
    doneSteps: int = 0
    failure: str = ''           # the device's failure of the run ('non-finite', 'time step collapsed')

    def __post_init__(self) -> None:
        if self.tracker is None:
            self.tracker = trackerOf(self.case)
        self.interval = trackInterval(self.case.parameters.get('depth', 1.0))
        self.nextTrack = self.time + self.interval


class Fleet:

    '''

    The runs of one dimension in one BatchEngine, stepped a tracking interval at a time.

    Parameters:
    -----------
    dimensions : int
        2 or 3
    precision : str
        'float32' or 'float64'
    lanes, block, listCapacity : int
        The launch shape, fixed for every engine the fleet builds so a run's sums do not depend on
        the mix
    chunk : int
        Steps per graph-replayed advance
    measure : str
        'device': the trackers' measures are taken on the device (pysurf_amd.readout) and only they
        come back; 'host': every reading downloads the state and the trackers measure it

    '''

    def __init__(self, dimensions: int, precision: str = 'float32', lanes: int = 2, block: int = 256,
                 listCapacity: int = -1, chunk: int = 8, device = None, measure: str = 'device') -> None:
        if measure not in ('device', 'host'):
            raise ValueError(f"A fleet measures on 'device' or 'host', not {measure!r}.")
        self.dimensions = dimensions
        self.precision = precision
        self.shape = (block, lanes, listCapacity)
        self.chunk = chunk
        self.device = device
        self.measure = measure
        self.readout = readout.DeviceReadout()
        self.members: list[Member] = []
        self.batch: BatchEngine | None = None
        self.updates = 0
        # device time of the fleet's steps (HIP events on its engine's stream around each interval's
        # launches), and the fluid updates those steps made
        self.deviceSeconds = 0.0
        self.fluidUpdates = 0
        self._interval = None

    def rebuild(self, leaving: list, joining: list) -> None:

        '''

        Drop `leaving` members and add `joining` ones, building the new engine from the members'
        current states: the members that stay are copied slot for slot from the old engine on the
        device (their state never crosses PCIe), the joining ones load from their cases.

        '''

        old = self.batch
        place = {id(member): index for index, member in enumerate(self.members)}
        keep = [m for m in self.members if m not in leaving]
        stash, starts = None, {}
        if old is not None:
            clock = old.clocks()
            for index, member in enumerate(self.members):
                member.time, member.steps = float(clock['time'][index]), int(clock['steps'][index])
            if keep:
                # the staying runs' state, gathered on the device into tensors just their size, so the
                # old engine (its neighbor lists and scratch) is freed before the new one is made
                old.synchronize()
                ranges = [(old.offsets[place[id(m)]], old.offsets[place[id(m)] + 1]) for m in keep]
                with torch.cuda.device(old.device):
                    stash = {name: torch.cat([old.engine.tensors[name][..., a:b] for a, b in ranges], dim = -1)
                             for name in stateFields}
                    torch.cuda.synchronize(old.device)
                start = 0
                for member, (a, b) in zip(keep, ranges):
                    starts[id(member)] = start
                    start += b - a
        self.members = keep + list(joining)
        self.batch = None
        del old
        if not self.members:
            return
        block, lanes, capacity = self.shape
        cases = [member.case for member in self.members]
        carry = [(stash, starts[id(member)]) if id(member) in starts else None for member in self.members]
        self.batch = BatchEngine(cases, self.precision, self.device, blockSize = block, lanes = lanes,
                                 listCapacity = capacity, carry = carry)
        for index, member in enumerate(self.members):
            self.batch.setClock(index, member.time, member.nextTrack, member.steps, 2**62)

    @property
    def particles(self) -> int:
        return sum(member.case.count for member in self.members)

    def launch(self) -> None:

        '''

        Point every member's clock at its next reading and enqueue enough graph-replayed steps to
        reach it at the last step sizes (a run's clock stops it at its reading, so surplus steps
        leave it alone); returns without waiting.

        '''

        clock = self.batch.clocks()
        needed = 1
        for index, member in enumerate(self.members):
            stopped = member.done or bool(member.failure)
            until = member.time if stopped else member.nextTrack
            self.batch.setClock(index, member.time, until, member.steps, 2**62)
            dt = float(clock['dt'][index])
            if not stopped and dt > 0.0:
                needed = max(needed, int(math.ceil((until - member.time) / dt)) + 1)
        stream = self.batch.engine.stream
        begin = torch.cuda.Event(enable_timing = True)
        begin.record(stream)
        self._interval = (begin, {index: member.steps for index, member in enumerate(self.members)})
        # the one graph of `chunk` steps, replayed: the engine keeps a single captured graph, so asking
        # for another step count would capture anew while the device waits
        for _ in range(max(1, -(-needed // self.chunk))):
            self.batch.advance(self.chunk)

    def collect(self) -> dict:

        '''

        Wait until every member's clock reaches its reading (stepping on where one falls short),
        then copy the positions back and advance each member's next reading; returns the reading.

        '''

        while not self.reached():
            self.batch.advance(self.chunk)
        if self._interval is not None:
            begin, before = self._interval
            end = torch.cuda.Event(enable_timing = True)
            end.record(self.batch.engine.stream)
            end.synchronize()
            self.deviceSeconds += begin.elapsed_time(end) / 1000.0
            self.fluidUpdates += sum(member.case.fluidCount * (member.steps - before[index])
                                     for index, member in enumerate(self.members))
            self._interval = None
        reading = {'offsets': list(self.batch.offsets), 'members': [], 'failed': []}
        for index, member in enumerate(self.members):
            if member.done:
                continue
            if member.failure:
                # a failed run is not measured: it ends at the tracking
                reading['failed'].append((member, member.time, member.steps))
                continue
            reading['members'].append((index, member, member.time, member.steps))
            member.nextTrack += member.interval
        entries = [readingEntry(index, member) for index, member, _, _ in reading['members']]
        if self.measure == 'device':
            reading['measures'] = self.readout.measure(self.batch, entries)
        else:
            reading['measures'] = readout.hostMeasures(self.batch, entries)
        return reading

    def reached(self) -> bool:
        '''Whether every member's clock has reached its reading or stopped on a failure; refreshes the
        members' clocks and failures.'''
        clock = self.batch.clocks()
        ready = True
        for index, member in enumerate(self.members):
            member.time, member.steps = float(clock['time'][index]), int(clock['steps'][index])
            code = int(clock['failed'][index])
            if code and not member.failure:
                member.failure = _native.failureNames[code]
            if not member.done and not member.failure and member.time < member.nextTrack:
                ready = False
        return ready

    def track(self, reading: dict, wall: float) -> list:
        '''Each member's tracker reads its measures of `reading`; returns the members that finished at
        it, a member whose run failed on the device among them.'''
        finished = []
        for member, time, steps in reading.get('failed', []):
            if member.done:
                continue
            member.done = True
            member.doneTime, member.doneSteps = time, steps
            member.finishedWall = wall
            finished.append(member)
        for index, member, time, steps in reading['members']:
            if member.done:
                continue
            measures = reading['measures'][index]
            if measures is not None:
                member.tracker.observeMeasures(time, measures)
            self.updates += 1
            if member.tracker.breakingFinal or time >= member.endTime:
                member.done = True
                member.doneTime, member.doneSteps = time, steps
                member.finishedWall = wall
                finished.append(member)
        return finished


def runToBreaking(cases: list, precision: str = 'float32', endTime: float = 10.0, lanes: int = 2,
                  listCapacity: int = -1, chunk: int = 8, measure: str = 'device') -> list:
    '''Step `cases` together to their breaking points; returns each run's tracker, in order.'''
    dimensions = {case.dimensions for case in cases}
    if len(dimensions) != 1:
        raise ValueError('One fleet steps runs of one dimension.')
    fleet = Fleet(dimensions.pop(), precision, lanes = lanes, listCapacity = listCapacity, chunk = chunk,
                  measure = measure)
    members = [Member(key = index, case = case, endTime = endTime) for index, case in enumerate(cases)]
    fleet.rebuild([], members)
    fleet.launch()
    while True:
        reading = fleet.collect()
        failed = [member for member in members if member.failure]
        if failed:
            # as the reference's run loop raises (BreakingWave.py:892-919)
            raise NumericalInstabilityError(', '.join(f'run {member.key} ({member.case.name}): {member.failure} at '
                                                      f't = {member.time:.6f} s' for member in failed))
        if not all(member.done for member in members):
            fleet.launch()
        fleet.track(reading, timeModule.perf_counter())
        if all(member.done for member in members):
            return [member.tracker for member in members]


@dataclass
class Request:

    '''A run to serve: arriving `arrival` seconds into the stream.'''

    key: int
    arrival: float
    case: object
    endTime: float = 10.0


@dataclass
class Served:

    '''What became of a request: arrival, admission and finish on the server's wall clock [s].'''

    key: int
    dimensions: int
    particles: int
    fluid: int
    arrival: float
    admitted: float
    finished: float
    simulated: float
    steps: int
    breaking: object
    tracker: object = None
    # why the request ended without its run finishing: 'budget' (its state can never fit the HBM
    # budget, refused on arrival), or the run's device failure ('non-finite', 'time step collapsed')
    refused: str = ''

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class Server:

    '''

    The request loop: requests arrive on the wall clock, wait until their state fits the HBM
    budget, join their dimension's fleet at the next interval boundary, and leave it when their
    breaking point is recorded or their time runs out. Both fleets' launches are enqueued before
    either is waited on, so their work overlaps on the device.

    '''

    budgetBytes: float = 200e9
    # the slots of each dimension: at most this many particles in flight in its fleets (0: no limit),
    # so the runs on the device finish at the pace one batch steps and the rest queue; a request
    # larger than its dimension's slots runs alone once that dimension's fleets are empty
    slots2D: int = 0
    slots3D: int = 0
    # 2D runs in a fleet per particle spacing: a fleet steps until its every member reaches its next
    # reading, and runs of one spacing take about as many steps to get there, so few stop and wait
    fleetsBySpacing: bool = False
    precision: str = 'float32'
    lanes: int = 4
    chunk: int = 8
    listCapacity2D: int = 64
    listCapacity3D: int = 768
    measure: str = 'device'
    fleets: dict = field(default_factory = dict)
    served: list = field(default_factory = list)
    queueDepth: list = field(default_factory = list)

    def footprint(self, case) -> float:
        capacity = self.listCapacity2D if case.dimensions == 2 else self.listCapacity3D
        return case.count * bytesPerParticle(case.dimensions, capacity)

    def fleetKey(self, case):
        '''The fleet a run joins: its dimension, or with fleetsBySpacing a 2D run's (2, spacing).'''
        if self.fleetsBySpacing and case.dimensions == 2:
            return (2, round(float(case.spacing), 9))
        return case.dimensions

    @staticmethod
    def dimensionOf(key) -> int:
        return key[0] if isinstance(key, tuple) else key

    def fleet(self, key) -> Fleet:
        if key not in self.fleets:
            dimensions = self.dimensionOf(key)
            capacity = self.listCapacity2D if dimensions == 2 else self.listCapacity3D
            self.fleets[key] = Fleet(dimensions, self.precision, lanes = self.lanes, listCapacity = capacity,
                                     chunk = self.chunk, measure = self.measure)
        return self.fleets[key]

    def refuse(self, request, now: float, reason: str) -> None:
        '''Record `request` as ended at `now` without a run, for `reason`.'''
        self.served.append(Served(key = request.key, dimensions = request.case.dimensions, particles = request.case.count,
                                  fluid = request.case.fluidCount, arrival = request.arrival, admitted = float('nan'),
                                  finished = now, simulated = 0.0, steps = 0, breaking = None, refused = reason))

    def serve(self, requests: list, wallLimit: float = float('inf')) -> list:

        '''

        Serve `requests` (sorted by arrival) until all finish or `wallLimit` seconds pass; returns
        the Served records. Each round launches every fleet's next interval, then reads the last
        interval's positions on the host while the device steps, so the trackers overlap the GPU.

        Admission takes the queue in arrival order and admits every request whose state fits the
        budget left and whose fleet has slots for its particles, so a large request waiting for room
        does not hold back smaller ones behind it (it keeps its place and is admitted as soon as room
        frees up). A request whose state alone exceeds the budget can never be admitted: it is
        refused on arrival.

        '''

        start = timeModule.perf_counter()
        clock = lambda: timeModule.perf_counter() - start
        waiting = list(requests)
        queued: list = []
        admittedAt: dict = {}
        inUse = 0.0
        pending = len(requests)
        readings: dict = {}
        while pending and clock() < wallLimit:
            now = clock()
            while waiting and waiting[0].arrival <= now:
                request = waiting.pop(0)
                if self.footprint(request.case) > self.budgetBytes:
                    self.refuse(request, now, 'budget')
                    pending -= 1
                else:
                    queued.append(request)
            # Membership changes while every fleet is idle: finished runs leave, admitted runs join
            joining: dict = {}
            still: list = []
            busy: dict = {}
            for key, fleet in self.fleets.items():
                d = self.dimensionOf(key)
                busy[d] = busy.get(d, 0) + sum(m.case.count for m in fleet.members if not m.done)
            for request in queued:
                d = request.case.dimensions
                slots = self.slots2D if d == 2 else self.slots3D
                fits = not slots or busy.get(d, 0) == 0 or busy.get(d, 0) + request.case.count <= slots
                if fits and inUse + self.footprint(request.case) <= self.budgetBytes:
                    inUse += self.footprint(request.case)
                    busy[d] = busy.get(d, 0) + request.case.count
                    admittedAt[request.key] = now
                    joining.setdefault(self.fleetKey(request.case), []).append(
                        Member(key = request, case = request.case, endTime = request.endTime))
                else:
                    still.append(request)
            queued = still
            self.queueDepth.append((now, len(queued)))
            for key in set(joining) | set(self.fleets):
                fleet = self.fleet(key)
                leaving = [member for member in fleet.members if member.done]
                if leaving or joining.get(key):
                    fleet.rebuild(leaving, joining.get(key, []))
            active = [fleet for fleet in self.fleets.values() if fleet.batch is not None
                      and not all(member.done for member in fleet.members)]
            if not active and not readings:
                # nothing on the device: wait for the next arrival (queued requests fit once room
                # frees, and with nothing running all the room is free, so none is left queued)
                if waiting:
                    timeModule.sleep(max(0.0, min(waiting[0].arrival - clock(), 0.05)))
                elif not queued:
                    break
                else:
                    timeModule.sleep(0.001)
                continue
            for fleet in active:
                fleet.launch()
            # The readings of the last round, tracked on the host while the device steps this one
            for fleet, reading in readings.items():
                for member in fleet.track(reading, clock()):
                    request = member.key
                    inUse -= self.footprint(request.case)
                    pending -= 1
                    self.served.append(Served(key = request.key, dimensions = request.case.dimensions,
                                              particles = request.case.count, fluid = request.case.fluidCount,
                                              arrival = request.arrival, admitted = admittedAt[request.key],
                                              finished = member.finishedWall, simulated = member.doneTime,
                                              steps = member.doneSteps,
                                              breaking = None if member.failure else member.tracker.breakingPoint,
                                              tracker = member.tracker, refused = member.failure))
            readings = {fleet: fleet.collect() for fleet in active}
            if not pending:
                break
        return self.served


def poissonRequests(cases: list, rate: float, seed: int = 0, endTime: float = 10.0) -> list:
    '''Requests for `cases` arriving as a Poisson stream of `rate` per second.'''
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, len(cases))
    arrivals = np.cumsum(gaps) - gaps[0]
    return [Request(key = index, arrival = float(arrival), case = case, endTime = endTime)
            for index, (arrival, case) in enumerate(zip(arrivals, cases))]


def latencyPercentiles(served: list) -> dict:
    '''p50 and p99 of the served runs' latencies (requests refused without a run left out).'''
    latencies = np.array([record.latency for record in served if not record.refused])
    if latencies.size == 0:
        return {'p50': float('nan'), 'p99': float('nan')}
    return {'p50': float(np.percentile(latencies, 50)), 'p99': float(np.percentile(latencies, 99))}
