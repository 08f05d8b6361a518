'''

Launch shapes for the pair sweeps, chosen by measurement on the device at hand.

A shape is (lanes per particle, threads per workgroup, neighbor-list entries per particle, and
optionally the cap in bytes of a sweep workgroup's LDS copy of the records around its particles and
the same cap for the step's list build apart).
More lanes per particle split each particle's neighbors over more lanes, which fills the GPU when a
context holds few particles and costs shuffles and idle lanes when it holds many; the list
capacity trades the list's HBM footprint against the particles that overflow it and walk their
cells; a larger LDS copy holds more of a dense 3D neighborhood but leaves room for fewer
workgroups. The best shape depends on the dimension, the pair precision, the particle count and
the smoothing length over the spacing, so the tuner times each candidate on the engine's own runs
and keeps the fastest.

Tuning steps the runs, so it reloads every run's initial state afterwards: a tuned engine starts
from the same state, with its clocks stopped, as a new one. Results are cached per (dimension,
precision, particle-count octave, smoothing lengths over spacings) for the life of the process.

'''

from __future__ import annotations

import math

# (lanes, workgroup threads, list capacity[, LDS copy cap[, the list build's cap]]); capacity -1 is
# 64 entries in 2D and 160 in 3D, a cap -1 32 KiB (the list build's: the sweeps'). A 2D workgroup's
# neighborhood fits 32 KiB at every lane count, but a smaller copy leaves room for one more workgroup
# on a CU; a 3D one outgrows it, so 3D also tries smaller copies for the same reason, for every
# kernel or for the list build alone.
CANDIDATES = {
    2: [(lanes, block, -1) for lanes in (1, 2, 4, 8, 16) for block in (128, 256)] +
       [(lanes, 256, -1, cap) for lanes in (4, 8) for cap in (12288, 16384)],
    3: [(lanes, block, -1) for lanes in (2, 4, 8, 16) for block in (128, 256)] +
       [(lanes, 256, -1, cap) for lanes in (4, 8) for cap in (0, 8192, 16384, 24576)] +
       [(lanes, 256, -1, -1, cap) for lanes in (4, 8) for cap in (8192, 16384)],
}

_cache: dict = {}


def cacheKey(dimensions: int, precision: str, particles: int, ratios: tuple = ()) -> tuple:
    return (dimensions, precision, int(math.log2(max(particles, 1))), ratios)


def lengthRatios(batch) -> tuple:
    '''The distinct smoothing lengths over spacings of the batch's runs, to two decimals.'''
    return tuple(sorted({round(case.smoothingLength / case.spacing, 2) for case in batch.cases}))


def candidatesFor(batch) -> list:

    '''

    The shapes to time on `batch`: CANDIDATES of its dimension, and where its smoothing lengths are
    long enough that the default list would overflow (a 2D particle has about pi (2 h / dx)^2
    neighbors, 21 at h = 1.3 dx but 113 at h = 3 dx), the same lanes with lists of 1.5 and 2.5 times
    that count, rounded up to 32 entries.

    '''

    shapes = list(CANDIDATES[batch.dimensions])
    ratio = max(lengthRatios(batch), default = 0.0)
    expected = math.pi * (2.0 * ratio)**2 if batch.dimensions == 2 else 4.0 / 3.0 * math.pi * (2.0 * ratio)**3
    default = 64 if batch.dimensions == 2 else 160
    if 1.25 * expected > default:
        for factor in (1.5, 2.5):
            capacity = min(4096, int(math.ceil(factor * expected / 32.0)) * 32)
            shapes += [(lanes, 256, capacity) for lanes in (4, 8, 16)]
    return shapes


def full(shape) -> tuple:
    '''A shape with its caps: (lanes, block, capacity, cap, the list build's cap).'''
    return tuple(shape) + (-1,) * (5 - len(shape))


def apply(batch, shape) -> None:
    '''Set a (lanes, block, capacity[, cap[, build cap]]) shape on the batch's engine.'''
    lanes, block, capacity, cap, build = full(shape)
    batch.engine.setLaunch(block, lanes, capacity)
    batch.engine.setNearCells(cap, build)


# Steps timed per candidate for a context of fewer particles than this, at the least: a small
# context's step is a fraction of a millisecond, and the timing needs several of them
SMALL_CONTEXT, SMALL_CONTEXT_STEPS = 1_000_000, 64


def timeSteps(batch, steps: int) -> float:
    '''Milliseconds per step of `steps` graph-replayed steps, between events on the engine's stream.'''
    import torch
    stream = batch.engine.stream
    # capture and warm the graph of `steps` steps first: a graph of another length would be captured
    # inside the timed interval, and for a small context the capture outlasts the steps
    batch.advance(steps, graph = True)
    batch.synchronize()
    start, end = torch.cuda.Event(enable_timing = True), torch.cuda.Event(enable_timing = True)
    start.record(stream)
    batch.advance(steps, graph = True)
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / steps


def tune(batch, steps: int = 8, candidates = None, reuse: bool = True) -> dict:

    '''

    Times every candidate shape on `batch` (a BatchEngine), sets the fastest and reloads the runs'
    initial state.

    Parameters:
    -----------
    batch : BatchEngine
        The engine to tune; its runs are stepped and then reloaded
    steps : int
        Timed steps per candidate
    candidates : list | None
        (lanes, block, listCapacity) shapes; None takes the dimension's defaults
    reuse : bool
        Take a cached result for the same dimension, precision and particle-count octave

    Returns:
    --------
    dict
        'lanes', 'block', 'listCapacity', 'ms_per_step' of the chosen shape and 'timings' of all

    '''

    key = cacheKey(batch.dimensions, batch.precision, batch.count, lengthRatios(batch))
    if reuse and key in _cache and candidates is None:
        chosen = _cache[key]
        apply(batch, (chosen['lanes'], chosen['block'], chosen['listCapacity'], chosen['nearCells'],
                      chosen['nearCellsBuild']))
        return chosen
    shapes = candidates if candidates is not None else candidatesFor(batch)
    if batch.count < SMALL_CONTEXT:
        steps = max(steps, SMALL_CONTEXT_STEPS)
    timings = []
    for shape in shapes:
        shape = full(shape)
        apply(batch, shape)
        batch.load()
        batch.setClocks(maxSteps = 2**62)
        timings.append((timeSteps(batch, steps),) + shape)
    best = min(timings)
    chosen = {'lanes': best[1], 'block': best[2], 'listCapacity': best[3], 'nearCells': best[4],
              'nearCellsBuild': best[5], 'ms_per_step': best[0],
              'timings': [{'lanes': l, 'block': b, 'listCapacity': c, 'nearCells': n, 'nearCellsBuild': nb,
                           'ms_per_step': t} for t, l, b, c, n, nb in timings]}
    apply(batch, (chosen['lanes'], chosen['block'], chosen['listCapacity'], chosen['nearCells'],
                  chosen['nearCellsBuild']))
    batch.load()
    if candidates is None:
        _cache[key] = chosen
    return chosen
