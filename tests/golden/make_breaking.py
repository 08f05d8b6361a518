'''

Writes breaking_c1_16ppd.json: the breaking point of the 16-particles-per-depth solitary wave on a
1:15 beach (H/d 0.45, crest at x = 4, seed 0) stepped by the float64 restatement (oracle/) and read
by the breaking tracker (pysurf_amd/breaking.py) every 0.01 sqrt(d / g), as the GPU test reads the
device's run. Run from the repository root: python tests/golden/make_breaking.py

'''

import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'tests'))

import numpy as np

from oracle import oracle
from pysurf_amd.breaking import trackInterval
from pysurf_amd.scenarios import solitaryOnSlope
from pysurf_amd.serving import trackerOf
from test_gpu_steps import physicsOf


def main() -> None:
    case = solitaryOnSlope(particlesPerDepth = 16)
    p = case.particles
    run = oracle.OracleRun(physicsOf(case), case.boundary.description(), p.positions, p.velocities, p.densities,
                           p.masses, p.isFluid)
    tracker = trackerOf(case)
    step = trackInterval(1.0)
    target = step
    while run.time.value < 10.0 and tracker.breakingPoint is None:
        run.advance(target, 10**7)
        tracker.update(run.time.value, run.positions[p.isFluid])
        target += step
    point = tracker.breakingPoint
    record = {'case': 'solitaryOnSlope(particlesPerDepth = 16), seed 0', 'interval': step, 'readings': len(tracker.times),
              'steps': run.steps, 'time': point.time, 'crestX': point.crestX, 'crestHeight': point.crestHeight,
              'depth': point.depth, 'lean': point.lean}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'breaking_c1_16ppd.json'), 'w') as f:
        json.dump(record, f, indent = 1)
    print(record)


if __name__ == '__main__':
    main()
