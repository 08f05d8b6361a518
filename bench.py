'''

Throughput of the MI355X delta-SPH step: fluid particle-updates per second on one GPU, batched
float32 pairs against the plain float64 kernel.

The headline workload is the batched sweep: 256 independent 2D runs (seeds 0..255: slope
U[1/35, 1/8], H/d U[0.2, 0.6], 16, 24, 32 or 48 particles per depth, length U[20, 30] d), each a
first-order solitary wave over a plane beach with the delta-SPH scheme (Wendland C2 at h = 1.3 dx,
c0 = 10 sqrt(g d), hydrostatic diffusion with delta = 0.1, extrapolated walls, L-corrected
continuity, limited viscosity with alpha = 0.01, density floor), packed into one context and
stepped by the same graph-replayed launches, with the launch shape chosen by the autotuner. A step
is one full device-resident step of every run: cell build, the three pair sweeps, kick, density
advance, drift and bed enforcement, each run with its own time step reduced on the device. Inputs
are resident in HBM before the timed region starts; one particle-update is one fluid particle
advanced one step.

Beside the headline the line carries:
  - the plain float64 kernel stepping a sample of the same runs, each alone, one after another;
  - the single 64-particles-per-depth run (float32 pairs and plain float64);
  - the 3D run: 12.4 M particles over a turned Gaussian ridge with a split crest zone;
  - the serving stream: Poisson arrivals, HBM admission, continuous batching, p50/p99 time to
    the breaking point;
  - Grilli's 1:35 case from the exact solitary wave at 64 particles per depth, stepped to the jet's
    touchdown and read by the smoothed-surface tracker, beside the reference's recorded run;
  - the dominant kernel's roofline, from HIP events on the engine's stream around each launch of a
    stream-launched pass over the same batch right after the timed region;
  - the float64 CPU restatement (oracle/) timed on the host's cores on a bounded sample of the
    sweep's runs.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

'''

from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = 'particle-updates/s on 1 MI355X, batched FP32 vs plain FP64; run p50/p99 latency'
HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of float32 kernels on the 256-run batched sweep (5,073,680 particles), keyed by
# (kernel, lanes per particle, particles): rocprofv3 FETCH_SIZE and WRITE_SIZE passes, each scaled by
# what the counter reads for a 1 GiB device copy (x 2.000 and x 1.000), as recorded in
# profiles/batched_sweep_256runs_counters.md. A kernel, shape or size without a measurement has none.
MEASURED_TRAFFIC = {
    ('buildNeighbors', 4, 5073680): 9.4210e+08,
    ('sweepContinuity', 4, 5073680): 1.3341e+09,
    ('sweepAccelerations', 4, 5073680): 1.0547e+09,
    ('gatherRecords', 4, 5073680): 7.3165e+08,
}


def kernelBytes(dimensions: int, real: int, entries: float) -> dict:

    '''

    Algorithmic bytes per particle of each kernel of the step: the particle's own inputs read once
    and its outputs written once, with `entries` neighbor-list entries per particle (its mean).
    Neighbor records read through the caches are not counted. Counted from the kernels in
    pysurf_amd/csrc/sph_hip.hip: a record is four quads of `real`-byte values.

    '''

    d, r, k = dimensions, real, entries
    record = 16 * r
    ids = 8                                      # run id and sorted order
    return {
        'cellKeys': 8 * d + 4 + 4 + 4,           # position; run id; key; the cell count's atomic
        'cellScatter': 4 + 4 + 4 + 4,            # key; cell start; the fill atomic; order
        'cellRank': 4 + 4 + 4,                   # the scattered slot; its key; the ordered slot
        'gatherRecords': ids + 16 * d + 8 + 2 * r + 1 + 4 + record + 4 + r,
        'buildNeighbors': ids + record + 4 + 4 + 4 * k + d * d * r,
        'sweepAccelerations': ids + record + 4 + d * d * r + 4 + 4 * k + 2 * d * r + d * d * r,
        'sweepContinuity': ids + record + 4 + d * d * r + 4 + 4 * k + 16 + 32 * d + 2 * d * r,
    }


# Useful float32 operations per neighbor-list entry of the 2D list build: the predicate on the pair
# that makes the entry (the separation from points and remainders 6, the squared distance 3, the
# compare 1) and sweep 1's sums of a fluid particle on it (the root 4, the Wendland slope over r 8,
# the volume 1, the velocity difference and the weighted separation 6, G's rises and offsets 16), as
# read from buildNeighbors and Sweep1::add in pysurf_amd/csrc/sph_hip.hip. Candidates that miss are
# not counted: their tests are the work the eighths tables exist to avoid.
FLOP_PER_ENTRY = {('buildNeighbors', 2): 51.0}
FP32_PEAK_TFLOPS = 157.3


def parse():
    parser = argparse.ArgumentParser(description = __doc__.split('\n\n')[1])
    parser.add_argument('--gpus', type = int, default = 1)
    parser.add_argument('--steps', type = int, default = 500, help = 'timed steps of the batched sweep')
    parser.add_argument('--warmup', type = int, default = 50)
    parser.add_argument('--chunk', type = int, default = 50, help = 'steps per captured graph')
    parser.add_argument('--runs', type = int, default = 256, help = 'runs of the batched sweep')
    parser.add_argument('--no-tune', action = 'store_true', help = 'the default launch shape instead of the tuned one')
    parser.add_argument('--lanes', type = int, default = 0, help = 'lanes per particle of the batched sweep (0: tuned)')
    parser.add_argument('--block', type = int, default = 256, help = 'workgroup threads with --lanes')
    parser.add_argument('--no-sort', action = 'store_true', help = 'keep the batched state in slot order')
    parser.add_argument('--profile-steps', type = int, default = 20, help = 'stream-launched steps timed per kernel')
    parser.add_argument('--plain-sample', type = int, default = 16, help = 'runs the plain float64 kernel steps alone')
    parser.add_argument('--plain-steps', type = int, default = 200)
    parser.add_argument('--single-steps', type = int, default = 1000, help = 'steps of the single 64 ppd run')
    parser.add_argument('--no-single', action = 'store_true', help = 'skip the single 64 ppd run')
    parser.add_argument('--ridge-steps', type = int, default = 100, help = 'timed steps of the 3D ridge run')
    parser.add_argument('--no-ridge', action = 'store_true', help = 'skip the 3D ridge run')
    parser.add_argument('--serving-window', type = float, default = 4.0, help = 'wall seconds of Poisson arrivals')
    parser.add_argument('--serving-pool', type = int, default = 40, help = 'distinct runs the requests draw from')
    parser.add_argument('--serving-wall', type = float, default = 120.0, help = 'wall seconds the serving stream may take')
    parser.add_argument('--serving-slots-2d', type = int, default = 4000000, help = '2D particles in flight')
    parser.add_argument('--serving-slots-3d', type = int, default = 3000000, help = '3D particles in flight')
    parser.add_argument('--serving-3d-end', type = float, default = 0.1, help = 'simulated seconds of a 3D request')
    parser.add_argument('--no-serving', action = 'store_true', help = 'skip the serving stream')
    parser.add_argument('--serving-one-fleet', action = 'store_true',
                        help = 'serve every 2D run in one fleet rather than a fleet per particle spacing')
    parser.add_argument('--no-gate', action = 'store_true', help = "skip Grilli's 1:35 case from the exact start")
    parser.add_argument('--gate-only', action = 'store_true',
                        help = "run only Grilli's case and print its record (for profiling the breaking read-out)")
    parser.add_argument('--single-only', action = 'store_true',
                        help = 'run only the single 64 ppd run and print its record (for profiling its launches)')
    parser.add_argument('--ridge-only', action = 'store_true',
                        help = 'run only the 3D ridge run and print its record (for profiling the 3D step)')
    parser.add_argument('--gate-end', type = float, default = 12.0,
                        help = "simulated seconds Grilli's case may run before its touchdown")
    parser.add_argument('--cpu-seconds', type = float, default = 15.0)
    parser.add_argument('--no-cpu', action = 'store_true', help = 'skip the CPU baseline')
    parser.add_argument('--copy-calibration', type = int, default = 0, metavar = 'MIB',
                        help = 'only copy a MIB-mebibyte device buffer 10 times (the counter calibration) and exit')
    return parser.parse_args()


def log(message: str) -> None:
    print(f'# {message}', file = sys.stderr, flush = True)


def timedSteps(batch, steps: int, chunk: int, sort: bool = False) -> tuple[float, float]:

    '''

    Wall seconds between device synchronizations, and event seconds on the engine's stream, for
    `steps` graph-replayed steps; with `sort`, the state is sorted into cell order after every chunk,
    inside the timed region.

    '''

    import torch
    stream = batch.engine.stream
    start, end = torch.cuda.Event(enable_timing = True), torch.cuda.Event(enable_timing = True)
    torch.cuda.synchronize()
    wall = time.perf_counter()
    start.record(stream)
    done = 0
    while done < steps:
        take = min(chunk, steps - done)
        batch.advance(take, graph = True)
        if sort:
            batch.sortParticles()
        done += take
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall
    return wall, start.elapsed_time(end) / 1000.0


def warm(batch, steps: int, chunk: int) -> None:
    batch.setClocks(maxSteps = 2**62)
    done = 0
    while done < steps:
        take = min(chunk, steps - done)
        batch.advance(take)
        done += take
    batch.synchronize()


def kernelProfile(batch, steps: int) -> dict:
    '''Per-kernel milliseconds and launches from HIP events on the engine's stream, over `steps` stream-launched steps.'''
    import ctypes

    import numpy as np

    from pysurf_amd import _native
    engine = batch.engine
    engine.check(engine.library.sph_profile(engine.handle, 1), 'sph_profile')
    batch.advance(steps, graph = False)
    kinds = len(_native.kernelNames)
    ms = np.zeros(kinds)
    launches = np.zeros(kinds, dtype = np.int64)
    engine.check(engine.library.sph_kernel_times(engine.handle, ms.ctypes.data_as(ctypes.c_void_p),
                                                 launches.ctypes.data_as(ctypes.c_void_p), kinds), 'sph_kernel_times')
    engine.check(engine.library.sph_profile(engine.handle, 0), 'sph_profile')
    return {name: {'ms': float(ms[k]), 'launches': int(launches[k])} for k, name in enumerate(_native.kernelNames)}


def meanNeighbors(batch) -> float:
    '''Mean neighbor count per particle of the batch's current state.'''
    import ctypes

    import torch
    engine = batch.engine
    counts = torch.zeros(max(batch.count, 1), dtype = torch.int32, device = batch.device)
    engine.check(engine.library.sph_neighbor_counts(engine.handle, ctypes.c_void_p(counts.data_ptr())),
                 'sph_neighbor_counts')
    return float(counts[:batch.count].double().mean().item())


def roofline(batch, profileSteps: int, precision: str) -> dict:
    '''The dominant kernel's algorithmic bytes over its mean launch time, from a profiled pass.'''
    profile = kernelProfile(batch, profileSteps)
    entries = meanNeighbors(batch)
    table = kernelBytes(batch.dimensions, 4 if precision == 'float32' else 8, entries)
    dominant = max((name for name in profile if profile[name]['launches'] and name in table),
                   key = lambda name: profile[name]['ms'])
    launchMs = profile[dominant]['ms'] / profile[dominant]['launches']
    perParticle = table[dominant]
    traffic = perParticle * batch.count
    achieved = traffic / (launchMs / 1000.0) / 1e9
    stepBytes = sum(table.values()) * batch.count
    stepMs = sum(entry['ms'] for entry in profile.values()) / profileSteps
    log(f'dominant kernel {dominant}: {launchMs * 1000:.1f} us/launch, {achieved:.1f} GB/s of algorithmic bytes')
    measured = MEASURED_TRAFFIC.get((dominant, batch.launch.get('lanes'), batch.count)) if precision == 'float32' else None
    flops = FLOP_PER_ENTRY.get((dominant, batch.dimensions)) if precision == 'float32' else None
    valu = None
    if flops:
        tflops = flops * entries * batch.count / (launchMs / 1000.0) / 1e12
        valu = {'flop_per_entry': flops, 'achieved': tflops, 'peak': FP32_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                'frac': tflops / FP32_PEAK_TFLOPS}
    return {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
            'traffic': measured, 'valu': valu, 'kernel': dominant, 'launch_us': launchMs * 1000.0,
            'algorithmic_bytes_per_launch': traffic, 'bytes_per_particle': perParticle,
            'mean_list_entries': entries, 'particles_per_launch': batch.count,
            'step_algorithmic_bytes_per_particle': sum(table.values()),
            'step_achieved_GBps': stepBytes / (stepMs / 1000.0) / 1e9, 'profiled_step_ms': stepMs,
            'kernel_us_per_step': {name: 1000.0 * entry['ms'] / profileSteps for name, entry in profile.items()
                                   if entry['launches']}}


def plainSample(cases, steps: int, chunk: int, tune: bool) -> dict:
    '''The plain float64 kernel stepping each run alone in its own context, one after another.'''
    from pysurf_amd.batch import BatchEngine
    updates, elapsed = 0, 0.0
    for case in cases:
        solo = BatchEngine([case], 'float64', tune = tune)
        warm(solo, chunk, chunk)
        elapsed += timedSteps(solo, steps, chunk)[0]
        updates += case.fluidCount * steps
        del solo
    value = updates / elapsed
    log(f'plain float64, {len(cases)} runs alone: {value:.4e} fluid updates/s')
    return {'value': value, 'unit': 'particle-updates/s', 'runs_sampled': len(cases), 'steps_each': steps,
            'how': 'each run alone in its own context, one after another, graph-replayed, tuned launch shape'}


def singleRun(args) -> dict:
    '''The single 64-particles-per-depth run: float32 pairs against the plain float64 kernel.'''
    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.scenarios import singleRun as single
    case = single(seed = 0, particlesPerDepth = 64)
    result = {'workload': f'{case.name}: H/d {case.parameters["waveHeight"]:.3f}, slope 1/'
                          f'{1.0 / case.parameters["slope"]:.1f}', 'particles': case.count,
              'fluid_particles': case.fluidCount, 'steps': args.single_steps}
    for precision in ('float32', 'float64'):
        batch = BatchEngine([case], precision, tune = not args.no_tune)
        warm(batch, args.warmup, args.chunk)
        wall, _ = timedSteps(batch, args.single_steps, args.chunk)
        profile = kernelProfile(batch, 20)
        result[precision] = {'value': case.fluidCount * args.single_steps / wall,
                             'ms_per_step': 1000.0 * wall / args.single_steps, 'lanes': batch.launch['lanes'],
                             'kernel_us_per_step': {name: 1000.0 * entry['ms'] / 20 for name, entry in profile.items()
                                                    if entry['launches']}}
        log(f'single run {precision}: {result[precision]["ms_per_step"]:.4f} ms/step')
        del batch
    result['speedup_of_f32'] = result['float32']['value'] / result['float64']['value']
    return result


def ridgeRun(args) -> dict:

    '''

    The 3D run: a solitary wave (H/d 0.5) over a Gaussian ridge turned 25 degrees in a 20 x 10 depth
    tank at 32 particles per depth, the crest zone split into daughters that keep h, float32 pairs,
    the launch shape tuned among neighbor lists long enough for the split zone.

    '''

    import numpy as np

    from pysurf_amd import autotune
    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.scenarios import ridge3D
    case = ridge3D()
    log(f'3D ridge: {case.count} particles ({case.fluidCount} fluid)')
    batch = BatchEngine([case], 'float32')
    launch = autotune.tune(batch, steps = 3, candidates = [(lanes, 256, 768) for lanes in (4, 8, 16)] +
                                                        [(8, 256, 768, cap) for cap in (0, 8192, 16384, 24576)] +
                                                        [(8, 256, 768, -1, cap) for cap in (8192, 16384)])
    before = batch.globalSums()[0]
    warm(batch, 2, args.chunk)
    # timed in segments of 100 steps (about 20 s), each reported as it ends
    wall, events, done = 0.0, 0.0, 0
    while done < args.ridge_steps:
        take = min(100, args.ridge_steps - done)
        segment = timedSteps(batch, take, args.chunk)
        wall, events, done = wall + segment[0], events + segment[1], done + take
        log(f'3D ridge: {done} of {args.ridge_steps} timed steps, {1000.0 * wall / done:.1f} ms/step')
    after = batch.globalSums()[0]
    clock = batch.clocks()
    profile = kernelProfile(batch, 2)
    entries = meanNeighbors(batch)
    value = case.fluidCount * args.ridge_steps / wall
    energy = lambda sums: float(sums[1] + sums[2] + sums[3])
    checks = {'steps': int(clock['steps'][0]), 'simulated_s': float(clock['time'][0]),
              'failed': int(clock['failed'][0]), 'finite': bool(np.all(np.isfinite(after))),
              'mass_relative_change': float(after[0] / before[0] - 1.0),
              'energy_relative_change': float(energy(after) / energy(before) - 1.0),
              'momentum_x_relative_change': float(after[4] / before[4] - 1.0)}
    log(f'3D ridge float32: {1000.0 * wall / args.ridge_steps:.1f} ms/step, {value:.4e} fluid updates/s')
    del batch
    return {'workload': f'{case.name}: Gaussian ridge 0.5 d tall, 2 d wide, 1:15, turned 25 degrees; 20 x 10 d tank; '
                        f'crest zone |x - 4 d| < 0.75 d split once, daughters keep h', 'particles': case.count,
            'fluid_particles': case.fluidCount, 'steps': args.ridge_steps, 'value': value,
            'unit': 'particle-updates/s', 'ms_per_step': 1000.0 * wall / args.ridge_steps,
            'mean_neighbors': entries,
            'launch': {key: launch[key] for key in ('lanes', 'block', 'listCapacity', 'nearCells', 'nearCellsBuild')},
            'checks': checks,
            'kernel_ms_per_step': {name: entry['ms'] / 2 for name, entry in profile.items() if entry['launches']}}


def servingRun(args) -> dict:

    '''

    The serving stream: requests arriving as a Poisson stream of 40 per second over a window of
    `args.serving_window` wall seconds, each a run drawn from a pool of `args.serving_pool` seeded
    runs (90% 2D runs of the batched sweep, 10% small 3D ridge runs at 16 particles per depth),
    admitted under a 200 GB HBM budget and the fleets' slots (`args.serving_slots_2d` and
    `_3d` particles in flight), and served by continuous batching until each run's breaking point
    is recorded (2D runs, at most 8 s simulated) or its time runs out (3D runs: `args.serving_3d_end`
    s simulated), for at most `args.serving_wall` wall seconds. The offered load is above one GPU's
    capacity, so requests queue for slots and the latencies include the queueing. Latency is each
    request's arrival to its breaking point's recording on the server's wall clock; requests
    unfinished at the wall limit are counted apart.

    '''

    import numpy as np

    from pysurf_amd.scenarios import servingMix
    from pysurf_amd.serving import Request, Server, latencyPercentiles
    pool = servingMix(args.serving_pool, seed = 0)
    kinds = {2: [case for case in pool if case.dimensions == 2], 3: [case for case in pool if case.dimensions == 3]}
    rng = np.random.default_rng(0)
    requests, arrival, rate = [], 0.0, 40.0
    while arrival < args.serving_window:
        dimensions = 3 if (rng.uniform() < 0.1 and kinds[3]) else 2
        case = kinds[dimensions][int(rng.integers(len(kinds[dimensions])))]
        requests.append(Request(key = len(requests), arrival = arrival, case = case,
                                endTime = args.serving_3d_end if dimensions == 3 else 8.0))
        arrival += float(rng.exponential(1.0 / rate))
    server = Server(budgetBytes = 200e9, slots2D = args.serving_slots_2d, slots3D = args.serving_slots_3d,
                    fleetsBySpacing = not args.serving_one_fleet)
    start = time.perf_counter()
    served = server.serve(requests, wallLimit = args.serving_wall)
    wall = time.perf_counter() - start
    finished = [record for record in served if not record.refused]
    broken = [record for record in finished if record.breaking is not None]
    updates = sum(record.fluid * record.steps for record in served)
    latency = latencyPercentiles(broken)
    depths = [depth for _, depth in server.queueDepth]
    log(f'serving: {len(requests)} requests over {args.serving_window:.0f} s, {len(finished)} finished in '
        f'{wall:.1f} s, {len(broken)} broke; time to breaking p50 {latency["p50"]:.2f} s, p99 {latency["p99"]:.2f} s; '
        f'queue depth max {max(depths, default = 0)}')
    return {'requests': len(requests), 'window_s': args.serving_window, 'pool': len(pool), 'finished': len(finished),
            'unfinished_at_wall_limit': len(requests) - len(served), 'refused': len(served) - len(finished),
            'broke': len(broken), 'rate_per_s': rate,
            'share_3d': float(np.mean([request.case.dimensions == 3 for request in requests])), 'budget_GB': 200.0,
            'wall_s': wall, 'wall_limit_s': args.serving_wall,
            'time_to_breaking_p50_s': latency['p50'], 'time_to_breaking_p99_s': latency['p99'],
            'completion_p50_s': latencyPercentiles(finished)['p50'],
            'completion_p99_s': latencyPercentiles(finished)['p99'],
            'throughput': updates / wall, 'unit': 'particle-updates/s',
            'slots_2d': args.serving_slots_2d, 'slots_3d': args.serving_slots_3d, 'end_3d_s': args.serving_3d_end,
            'fleets_by_spacing': server.fleetsBySpacing,
            'fleets': sorted(({'fleet': str(key), 'fluid_updates': fleet.fluidUpdates, 'device_s': fleet.deviceSeconds,
                               'updates_per_device_s': fleet.fluidUpdates / fleet.deviceSeconds
                               if fleet.deviceSeconds > 0 else None}
                              for key, fleet in server.fleets.items()), key = lambda entry: entry['fleet']),
            'max_queue_depth': max(depths, default = 0),
            'mean_queue_depth': float(np.mean(depths)) if depths else 0.0,
            'particles_max': max(case.count for case in pool), 'particles_min': min(case.count for case in pool)}


def gateRun(args) -> dict:

    '''

    Grilli's 1:35 case from the exact solitary wave at 64 particles per depth (128,947 fluid
    particles, Wendland kernel at h = 3 spacings, bed-fitted columns), float32 pairs, read by the
    smoothed-surface tracker every 0.01 sqrt(d/g) until the jet's touchdown fixes the breaking
    point; reported beside the reference's own GPU run of the same case
    (validation/gpuSph/results/grilliBreaking_64.json: 693 s on an RTX 2070 Super).

    '''

    from pysurf_amd.scenarios import exactSolitaryOnSlope
    from pysurf_amd.serving import Fleet, Member
    case = exactSolitaryOnSlope(64)
    fleet = Fleet(2, 'float32', lanes = 4, listCapacity = 160, chunk = 8)
    member = Member(key = 0, case = case, endTime = args.gate_end)
    start = time.perf_counter()
    fleet.rebuild([], [member])
    fleet.launch()
    readings = 0
    while not member.done:
        reading = fleet.collect()
        if not member.done:
            fleet.launch()
        fleet.track(reading, time.perf_counter())
        readings += 1
        if readings % 1000 == 0:
            log(f'gate: {readings} readings, t = {member.time:.2f} s')
    wall = time.perf_counter() - start
    tracker, point = member.tracker, member.tracker.breakingPoint
    result = {'case': 'exact solitary wave H/d = 0.3 on a 1:35 beach, 64 ppd', 'fluid_particles': case.fluidCount,
              'steps': member.steps, 'readings': readings, 'wall_s': wall, 'touchdown_s': tracker.touchdownTime,
              'breaking_s': point.time if point else None, 'Hb_over_h0': point.crestHeight if point else None,
              'hb_over_h0': point.depth if point else None,
              'reference_record': {'breaking_s': 8.7945, 'Hb_over_h0': 0.4963, 'hb_over_h0': 0.3159,
                                   'touchdown_s': 9.3740, 'wall_s': 693.1},
              'grilli_1997': {'Hb_over_h0': 0.476, 'hb_over_h0': 0.345},
              'fluid_updates_per_s': case.fluidCount * member.steps / wall}
    if point is not None:
        log(f'gate: breaking at {point.time:.4f} s, Hb/h0 {point.crestHeight:.4f}, hb/h0 {point.depth:.4f}, '
            f'touchdown {tracker.touchdownTime:.4f} s; {member.steps} steps in {wall:.1f} s')
    else:
        log(f'gate: no breaking point by {member.time:.3f} s; {member.steps} steps in {wall:.1f} s')
    return result


def cpuModel() -> str:
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def oracleRunOf(case):
    '''The float64 restatement (oracle/sph_oracle.cpp) of `case` from its start.'''
    from oracle import oracle
    p = case.particles
    physics = oracle.Physics(dims = case.dimensions, h = case.smoothingLength, domainMin = case.domainMin,
                             domainMax = case.domainMax, kernelType = case.kernelType, rho0 = case.referenceDensity,
                             c0 = case.soundSpeed, alpha = case.alpha, boundaryFactor = case.boundaryViscosityFactor,
                             gravity = case.gravity, diffusion = case.diffusion, delta = case.delta,
                             extrapolatedWalls = case.extrapolatedWalls, correction = case.correction,
                             limiter = case.limiter, densityFloor = case.densityFloor, cfl = case.cfl,
                             maxTimeStep = case.maxTimeStep)
    return oracle.OracleRun(physics, case.boundary.description(), p.positions, p.velocities, p.densities, p.masses,
                            p.isFluid)


def cpuBaseline(seconds: float) -> dict:

    '''

    The float64 CPU restatement (oracle/sph_oracle.cpp) on the host's cores, on the reference's
    CPU-runnable case and a subset of the single-run case: the 16 particles-per-depth solitary wave
    (H/d 0.45, 1:15 beach) stepped its 3000 steps from the start, then the 64 particles-per-depth
    run (H/d and slope from seed 0) for as many steps as fill about a sixth of `seconds`. Three
    repeats after a warm-up, each from the runs' starts; the median of their fluid updates per second.

    '''

    from oracle import oracle
    from pysurf_amd.scenarios import singleRun, solitaryOnSlope
    threads = min(16, os.cpu_count() or 1)
    oracle.threads(threads)
    small, single = solitaryOnSlope(particlesPerDepth = 16), singleRun(seed = 0, particlesPerDepth = 64)
    warm = oracleRunOf(single)
    start = time.perf_counter()
    warm.advance(float('inf'), 2)
    perStep = max((time.perf_counter() - start) / 2.0, 1e-6)
    subset = max(5, int(seconds / 6.0 / perStep))
    rates = []
    for _ in range(3):
        runs = [(oracleRunOf(small), 3000, small.fluidCount), (oracleRunOf(single), subset, single.fluidCount)]
        start = time.perf_counter()
        for run, steps, _ in runs:
            assert run.advance(float('inf'), steps) == steps
        elapsed = time.perf_counter() - start
        rates.append(sum(steps * fluid for _, steps, fluid in runs) / elapsed)
    value = statistics.median(rates)
    log(f'cpu: {value:.4e} fluid updates/s')
    return {'value': value, 'unit': 'particle-updates/s', 'cores': threads, 'kind': 'port',
            'sample': f'the float64 C++ OpenMP restatement (oracle/sph_oracle.cpp) stepping {small.name} '
                      f'({small.count} particles) its 3000 steps and {single.name} ({single.count} particles) '
                      f'{subset} steps, each from its start, median of 3 repeats',
            'repeats': rates, 'nproc': os.cpu_count(), 'cpu_model': cpuModel()}


def copyCalibration(mebibytes: int) -> int:
    '''Copies a device buffer of `mebibytes` MiB ten times: the known-size kernel the HBM counters are scaled by.'''
    import torch
    count = mebibytes * (1 << 20) // 4
    source = torch.ones(count, dtype = torch.float32, device = 'cuda')
    target = torch.empty_like(source)
    for _ in range(10):
        target.copy_(source)
    torch.cuda.synchronize()
    print(json.dumps({'copy_bytes_read': count * 4, 'copy_bytes_written': count * 4, 'copies': 10}))
    return 0


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f'bench.py measures one GPU: every run here fits one MI355X and nothing is sharded across GPUs '
              f'(asked for --gpus {args.gpus}).', file = sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print('bench.py needs a HIP device.', file = sys.stderr)
        return 1
    if args.copy_calibration:
        return copyCalibration(args.copy_calibration)
    if args.gate_only:
        print(json.dumps(gateRun(args)))
        return 0
    if args.single_only:
        print(json.dumps(singleRun(args)))
        return 0
    if args.ridge_only:
        print(json.dumps(ridgeRun(args)))
        return 0

    import numpy as np

    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.scenarios import sweepRun
    cases = [sweepRun(seed) for seed in range(args.runs)]
    fluid = sum(case.fluidCount for case in cases)
    particles = sum(case.count for case in cases)
    log(f'batched sweep: {len(cases)} runs, {particles} particles ({fluid} fluid)')
    if args.lanes:
        batch = BatchEngine(cases, 'float32', blockSize = args.block, lanes = args.lanes)
    else:
        batch = BatchEngine(cases, 'float32', tune = not args.no_tune)
    log(f'launch shape: {batch.launch.get("lanes")} lanes, {batch.launch.get("block")} threads')
    warm(batch, args.warmup, args.chunk)
    if not args.no_sort:
        batch.sortParticles()
    wall, events = timedSteps(batch, args.steps, args.chunk, sort = not args.no_sort)
    value = fluid * args.steps / wall
    msPerStep = 1000.0 * wall / args.steps
    clock = batch.clocks()
    log(f'batched float32: {msPerStep:.3f} ms/step, {value:.4e} fluid updates/s')
    roof = roofline(batch, args.profile_steps, 'float32')
    launch = dict(batch.launch)
    del batch

    plain = plainSample(cases[:args.plain_sample], args.plain_steps, args.chunk, not args.no_tune) \
        if args.plain_sample > 0 else None
    single = None if args.no_single else singleRun(args)
    ridge = None if args.no_ridge else ridgeRun(args)
    serving = None if args.no_serving else servingRun(args)
    gate = None if args.no_gate else gateRun(args)
    cpu = None if args.no_cpu else cpuBaseline(args.cpu_seconds)

    sizes = np.array([case.count for case in cases])
    line = {
        'metric': METRIC,
        'value': value,
        'unit': 'particle-updates/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': msPerStep,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic: first-order solitary waves on plane beaches laid on the lattice with 1e-3 dx jitter, '
                'drawn per seed as the batched-sweep configuration states',
        'config': {'workload': f'batched sweep of {len(cases)} 2D delta-SPH solitary-wave runs (seeds 0..'
                               f'{len(cases) - 1}: slope U[1/35, 1/8], H/d U[0.2, 0.6], 16/24/32/48 particles per '
                               f'depth, length U[20, 30] d) in one context, float32 pairs, float64 state',
                   'runs': len(cases), 'particles': particles, 'fluid_particles': fluid,
                   'run_particles_min': int(sizes.min()), 'run_particles_max': int(sizes.max()),
                   'kernel': 'Wendland C2, h = 1.3 dx', 'graph_chunk': args.chunk,
                   'state_sort': None if args.no_sort else f'into cell order after every {args.chunk} steps, timed',
                   'launch': launch,
                   'event_ms_per_step': 1000.0 * events / args.steps,
                   'simulated_seconds_min': float(clock['time'].min()),
                   'simulated_seconds_max': float(clock['time'].max())},
        'roofline': roof,
        'plain_fp64': plain,
        'speedup_over_plain_fp64': value / plain['value'] if plain else None,
        'single_run': single,
        'ridge_3d': ridge,
        'serving': serving,
        'grilli_gate': gate,
        'cpu_baseline': cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == '__main__':
    sys.exit(main())
