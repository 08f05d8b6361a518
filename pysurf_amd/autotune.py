'''

Launch shapes for the pair sweeps, chosen by measurement on the device at hand.

A shape is (lanes per particle, threads per workgroup, neighbor-list entries per particle). More
lanes per particle split each particle's neighbors over more lanes, which fills the GPU when a
context holds few particles and costs shuffles and idle lanes when it holds many; the list
capacity trades the list's HBM footprint against the particles that overflow it and walk their
cells. The best shape depends on the dimension, the pair precision and the particle count, so the
tuner times each candidate on the engine's own runs and keeps the fastest.

Tuning steps the runs, so it reloads every run's initial state afterwards: a tuned engine starts
from the same state, with its clocks stopped, as a new one. Results are cached per (dimension,
precision, particle-count octave) for the life of the process.

'''

from __future__ import annotations

import math

# (lanes, workgroup threads, list capacity); capacity -1 is 64 entries in 2D and 160 in 3D
CANDIDATES = {
    2: [(lanes, block, -1) for lanes in (1, 2, 4, 8, 16) for block in (128, 256)],
    3: [(lanes, block, -1) for lanes in (2, 4, 8, 16) for block in (128, 256)],
}

_cache: dict = {}


def cacheKey(dimensions: int, precision: str, particles: int) -> tuple:
    return (dimensions, precision, int(math.log2(max(particles, 1))))


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

    key = cacheKey(batch.dimensions, batch.precision, batch.count)
    if reuse and key in _cache and candidates is None:
        chosen = _cache[key]
        batch.engine.setLaunch(chosen['block'], chosen['lanes'], chosen['listCapacity'])
        return chosen
    shapes = candidates if candidates is not None else CANDIDATES[batch.dimensions]
    if batch.count < SMALL_CONTEXT:
        steps = max(steps, SMALL_CONTEXT_STEPS)
    timings = []
    for lanes, block, capacity in shapes:
        batch.engine.setLaunch(block, lanes, capacity)
        batch.load()
        batch.setClocks(maxSteps = 2**62)
        timings.append((timeSteps(batch, steps), lanes, block, capacity))
    best = min(timings)
    chosen = {'lanes': best[1], 'block': best[2], 'listCapacity': best[3], 'ms_per_step': best[0],
              'timings': [{'lanes': l, 'block': b, 'listCapacity': c, 'ms_per_step': t} for t, l, b, c in timings]}
    batch.engine.setLaunch(chosen['block'], chosen['lanes'], chosen['listCapacity'])
    batch.load()
    if candidates is None:
        _cache[key] = chosen
    return chosen
