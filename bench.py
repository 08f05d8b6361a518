'''

Throughput of the MI355X delta-SPH step: fluid particle-updates per second on one GPU.

The workload is the single 2D run of the benchmark's second configuration: a first-order solitary
wave over a plane beach at 64 particles per depth, H/d and slope drawn from seed 0, Wendland C2 at
h = 1.3 dx, c0 = 10 sqrt(g d), the delta-SPH scheme (hydrostatic diffusion with delta = 0.1,
extrapolated walls, L-corrected continuity, limited viscosity with alpha = 0.01, density floor),
synthetic input laid as the reference lays it. A step is one full device-resident step: cell build,
the three pair sweeps, kick, density advance, drift and bed enforcement, with the time step reduced
on the device. Steps are replayed as HIP graphs of `--chunk` steps; inputs are resident in HBM
before the timed region starts.

Beside the headline the line carries the plain float64 kernel's rate on the same run, the
dominant kernel's roofline from HIP events recorded around its launches during a profiled pass,
and the float64 CPU restatement timed on the host's cores on a bounded sample of the same run.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

'''

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = 'particle-updates/s on 1 MI355X, batched FP32 vs plain FP64; run p50/p99 latency'
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle of each kernel: the particle's own inputs read once and its outputs
# written once (neighbors' records are reused from cache and not counted), by dimension and pair
# precision. Counted from the kernels in pysurf_amd/csrc/sph_hip.hip.
def kernelBytes(dimensions: int, real: int) -> dict:
    d, r = dimensions, real
    lo = d * 4 if r == 4 else 0                                        # float32 remainders
    record = d * r + lo + d * r + 5 * r + 1 + 4                        # pos, lo, vel, rho, press, dev, mass, len, fluid, key
    index = 8                                                          # run id and sorted order
    return {
        'gatherRecords': index + d * 8 * 2 + 8 + 2 * r + 1 + 4 + record + r,
        'sweepWallAndGradients': index + 1 + d * r + lo + r + 4 + d * r + 2 * r + d * d * r,
        'sweepAccelerations': index + 1 + d * r + lo + r + 4 + d * r + 3 * r + d * d * r + d * r + d * d * r,
        'kick': index + 1 + d * r + 2 * d * 8 + d * r,
        'sweepContinuity': index + 1 + d * r + lo + r + 4 + d * r + r + d * d * r + 2 * 8 + 4 * d * 8,
        'buildNeighbors': index + d * r + lo + r + 4 + 4,
    }


def parse():
    parser = argparse.ArgumentParser(description = __doc__.split('\n\n')[1])
    parser.add_argument('--gpus', type = int, default = 1)
    parser.add_argument('--steps', type = int, default = 2000)
    parser.add_argument('--warmup', type = int, default = 100)
    parser.add_argument('--chunk', type = int, default = 50, help = 'steps per captured graph')
    parser.add_argument('--fp64-steps', type = int, default = 400)
    parser.add_argument('--profile-steps', type = int, default = 50)
    parser.add_argument('--cpu-seconds', type = float, default = 15.0)
    parser.add_argument('--seed', type = int, default = 0)
    parser.add_argument('--no-cpu', action = 'store_true', help = 'skip the CPU baseline')
    parser.add_argument('--no-batched', action = 'store_true', help = 'skip the batched sweep')
    parser.add_argument('--runs', type = int, default = 256, help = 'runs of the batched sweep')
    parser.add_argument('--batched-steps', type = int, default = 200)
    parser.add_argument('--plain-sample', type = int, default = 16, help = 'runs the plain float64 kernel steps alone')
    parser.add_argument('--plain-steps', type = int, default = 200)
    return parser.parse_args()


def timedSteps(batch, steps: int, chunk: int) -> float:
    '''Seconds for `steps` graph-replayed steps, between events on the engine's stream.'''
    import torch
    stream = batch.engine.stream
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing = True), torch.cuda.Event(enable_timing = True)
    start.record(stream)
    done = 0
    while done < steps:
        take = min(chunk, steps - done)
        batch.advance(take, graph = True)
        done += take
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / 1000.0


def prepare(case, precision: str, steps: int, chunk: int, warmup: int):
    from pysurf_amd.batch import BatchEngine
    batch = BatchEngine([case], precision)
    batch.setClocks(maxSteps = 2**62)
    done = 0
    while done < warmup:
        take = min(chunk, warmup - done)
        batch.advance(take)
        done += take
    batch.synchronize()
    return batch


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


def batchedSweep(args) -> dict:

    '''

    The batched sweep: `args.runs` uneven 2D runs (slope, height, resolution and length drawn per
    seed) packed into one context and stepped by the same graph-replayed launches, against the plain
    float64 kernel stepping a sample of the same runs, each alone, one after another.

    '''

    import numpy as np

    from pysurf_amd.batch import BatchEngine
    from pysurf_amd.scenarios import sweepRun
    cases = [sweepRun(seed) for seed in range(args.runs)]
    fluid = sum(case.fluidCount for case in cases)
    particles = sum(case.count for case in cases)
    print(f'# batched: {len(cases)} runs, {particles} particles ({fluid} fluid)', file = sys.stderr, flush = True)
    batch = BatchEngine(cases, 'float32')
    batch.setClocks(maxSteps = 2**62)
    batch.advance(args.chunk)
    batch.synchronize()
    seconds = timedSteps(batch, args.batched_steps, args.chunk)
    value = fluid * args.batched_steps / seconds
    clock = batch.clocks()
    del batch
    print(f'# batched float32: {1000 * seconds / args.batched_steps:.3f} ms/step, {value:.4e} fluid updates/s',
          file = sys.stderr, flush = True)
    updates, elapsed = 0, 0.0
    sample = cases[:args.plain_sample]
    for case in sample:
        solo = BatchEngine([case], 'float64')
        solo.setClocks(maxSteps = 2**62)
        solo.advance(args.chunk)
        solo.synchronize()
        elapsed += timedSteps(solo, args.plain_steps, args.chunk)
        updates += case.fluidCount * args.plain_steps
        del solo
    plain = updates / elapsed
    print(f'# plain float64, {len(sample)} runs alone: {plain:.4e} fluid updates/s', file = sys.stderr, flush = True)
    sizes = np.array([case.count for case in cases])
    return {'value': value, 'unit': 'particle-updates/s', 'runs': len(cases), 'particles': particles,
            'fluid_particles': fluid, 'run_particles_min': int(sizes.min()), 'run_particles_max': int(sizes.max()),
            'steps': args.batched_steps, 'ms_per_step': 1000.0 * seconds / args.batched_steps,
            'simulated_seconds_min': float(clock['time'].min()), 'simulated_seconds_max': float(clock['time'].max()),
            'plain_fp64': {'value': plain, 'unit': 'particle-updates/s', 'runs_sampled': len(sample),
                           'steps_each': args.plain_steps, 'how': 'each run alone in its own context, one after another'},
            'speedup_over_plain_fp64': value / plain,
            'workload': 'batched sweep: runs seeded 0..N-1 with slope U[1/35, 1/8], H/d U[0.2, 0.6], '
                        'resolution from {16, 24, 32, 48} particles per depth, length U[20, 30] d, float32 pairs'}


def cpuBaseline(case, seconds: float) -> dict:
    '''The float64 CPU restatement on the host's cores, over as many steps of the same run as fit `seconds`.'''
    import numpy as np

    from oracle import oracle
    threads = min(16, os.cpu_count() or 1)
    oracle.threads(threads)
    p = case.particles
    physics = oracle.Physics(dims = case.dimensions, h = case.smoothingLength, domainMin = case.domainMin,
                             domainMax = case.domainMax, kernelType = case.kernelType, rho0 = case.referenceDensity,
                             c0 = case.soundSpeed, alpha = case.alpha, boundaryFactor = case.boundaryViscosityFactor,
                             gravity = case.gravity, diffusion = case.diffusion, delta = case.delta,
                             extrapolatedWalls = case.extrapolatedWalls, correction = case.correction,
                             limiter = case.limiter, densityFloor = case.densityFloor)
    run = oracle.OracleRun(physics, case.boundary.description(), p.positions, p.velocities, p.densities, p.masses,
                           p.isFluid)
    run.advance(np.inf, 1)                      # warm-up step
    steps, start = 0, time.perf_counter()
    while time.perf_counter() - start < seconds or steps < 2:
        run.advance(np.inf, 1)
        steps += 1
    elapsed = time.perf_counter() - start
    return {'value': case.fluidCount * steps / elapsed, 'unit': 'particle-updates/s', 'cores': threads, 'kind': 'port',
            'sample': f'{steps} steps of the same {case.count}-particle run in the float64 C++ OpenMP restatement '
                      f'(oracle/sph_oracle.cpp), {elapsed:.1f} s, {threads} threads'}


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f'bench.py measures one GPU: every run here fits one MI355X and nothing is sharded across GPUs '
              f'(asked for --gpus {args.gpus}).', file = sys.stderr)
        return 2
    import torch

    from pysurf_amd.scenarios import singleRun
    if not torch.cuda.is_available():
        print('bench.py needs a HIP device.', file = sys.stderr)
        return 1
    case = singleRun(seed = args.seed, particlesPerDepth = 64)
    print(f'# workload {case.name}: {case.count} particles ({case.fluidCount} fluid)', file = sys.stderr, flush = True)

    batch = prepare(case, 'float32', args.steps, args.chunk, args.warmup)
    print('# float32 warm-up done', file = sys.stderr, flush = True)
    seconds = timedSteps(batch, args.steps, args.chunk)
    value = case.fluidCount * args.steps / seconds
    msPerStep = 1000.0 * seconds / args.steps
    print(f'# float32: {msPerStep:.4f} ms/step, {value:.4e} fluid updates/s', file = sys.stderr, flush = True)

    profile = kernelProfile(batch, args.profile_steps)
    table = kernelBytes(case.dimensions, 4)
    dominant = max((name for name in profile if profile[name]['launches'] and name in table),
                   key = lambda name: profile[name]['ms'])
    launchMs = profile[dominant]['ms'] / profile[dominant]['launches']
    perParticle = table[dominant]
    traffic = perParticle * case.count
    achieved = traffic / (launchMs / 1000.0) / 1e9
    stepMs = sum(entry['ms'] for entry in profile.values()) / args.profile_steps
    print(f'# dominant kernel {dominant}: {launchMs * 1000:.2f} us/launch, {achieved:.1f} GB/s', file = sys.stderr,
          flush = True)
    del batch

    plain = prepare(case, 'float64', args.fp64_steps, args.chunk, min(args.warmup, 50))
    fp64Seconds = timedSteps(plain, args.fp64_steps, args.chunk)
    fp64Value = case.fluidCount * args.fp64_steps / fp64Seconds
    print(f'# float64: {1000.0 * fp64Seconds / args.fp64_steps:.4f} ms/step, {fp64Value:.4e} fluid updates/s',
          file = sys.stderr, flush = True)
    del plain

    batched = None
    if not args.no_batched:
        batched = batchedSweep(args)

    cpu = None
    if not args.no_cpu:
        cpu = cpuBaseline(case, args.cpu_seconds)
        print(f'# cpu: {cpu["value"]:.4e} fluid updates/s ({cpu["sample"]})', file = sys.stderr, flush = True)

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
        'data': 'synthetic: first-order solitary wave on a plane beach laid on the lattice with 1e-3 dx jitter, '
                'seeded as the benchmark configuration states',
        'config': {'workload': f'2D delta-SPH solitary wave on a slope, 64 particles per depth, single run '
                               f'({case.name}: H/d {case.parameters["waveHeight"]:.3f}, slope 1/'
                               f'{1.0 / case.parameters["slope"]:.1f})',
                   'particles': case.count, 'fluid_particles': case.fluidCount, 'state': 'f64 positions, '
                   'velocities, densities; f32 pair sums', 'kernel': 'Wendland C2, h = 1.3 dx',
                   'graph_chunk': args.chunk},
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK_GBS, 'traffic': None, 'kernel': dominant,
                     'launch_us': launchMs * 1000.0, 'algorithmic_bytes_per_launch': traffic,
                     'bytes_per_particle': perParticle, 'profiled_step_ms': stepMs,
                     'kernel_ms_per_step': {name: entry['ms'] / args.profile_steps for name, entry in profile.items()}},
        'plain_fp64': {'value': fp64Value, 'unit': 'particle-updates/s', 'steps': args.fp64_steps,
                       'ms_per_step': 1000.0 * fp64Seconds / args.fp64_steps, 'speedup_of_f32': value / fp64Value},
        'batched': batched,
        'cpu_baseline': cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == '__main__':
    sys.exit(main())
