'''

Writes crest_c1_16ppd.json: the crest's track of the 16-particles-per-depth solitary wave on a 1:15
beach (H/d 0.45, crest at x = 4, seed 0) over its first 3 s, stepped by the float64 restatement
(oracle/) and read by the breaking tracker (pysurf_amd/breaking.py) every 0.01 sqrt(d / g), as the
GPU test reads the device's run: each reading's time, crest position and crest height. Run from the
repository root: python tests/golden/make_crest.py

'''

import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'tests'))

from oracle import oracle
from pysurf_amd.breaking import trackInterval
from pysurf_amd.scenarios import solitaryOnSlope
from pysurf_amd.serving import trackerOf
from test_gpu_steps import physicsOf

END = 3.0


def main() -> None:
    case = solitaryOnSlope(particlesPerDepth = 16)
    p = case.particles
    run = oracle.OracleRun(physicsOf(case), case.boundary.description(), p.positions, p.velocities, p.densities,
                           p.masses, p.isFluid)
    tracker = trackerOf(case)
    step = trackInterval(1.0)
    target = step
    while target <= END:
        run.advance(target, 10**7)
        tracker.update(run.time.value, run.positions[p.isFluid])
        target += step
    record = {'case': 'solitaryOnSlope(particlesPerDepth = 16), seed 0', 'interval': step, 'steps': run.steps,
              'time': tracker.times, 'crestX': tracker.crestX, 'crestHeight': tracker.crestHeight}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'crest_c1_16ppd.json'), 'w') as f:
        json.dump(record, f)
    print(len(tracker.times), 'readings,', run.steps, 'steps')


if __name__ == '__main__':
    main()
