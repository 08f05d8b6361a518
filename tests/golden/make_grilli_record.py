'''

Writes tests/golden/grilli_breaking_64.json and grilli_breaking_128_antuono.json: the breaking
points the reference recorded for Grilli's 1:35 case at 64 particles per depth, and at 128 with the
renormalized 'antuono' diffusion, on its GPU backend, copied from its results files
validation/gpuSph/results/grilliBreaking_64.json and grilliBreaking_128_antuono.json (data, read as
JSON; no reference code runs), with Grilli, Svendsen and Subramanya's (1997) values they are judged
against.

    python tests/golden/make_grilli_record.py [reference root]

'''

import json
import os
import sys

FIELDS = ('particlesPerDepth', 'bedFitted', 'faceArc', 'fluidParticles', 'breakingTime', 'breakingHeightOverDepth',
          'breakingDepthOverDepth', 'breakingFaceDepthOverDepth', 'breakingTroughOverDepth', 'touchdownTime',
          'touchedDown')


def write(root: str, run: str, case: str, target: str) -> None:
    source = os.path.join(root, 'validation', 'gpuSph', 'results', run + '.json')
    record = json.load(open(source))
    summary = record['summary']
    out = {'source': f'validation/gpuSph/results/{run}.json (the reference GPU backend run of '
                     f'validation/gpuSph/grilliBreaking.py {summary["particlesPerDepth"]})',
           'case': case,
           'recorded': {name: summary[name] for name in FIELDS},
           'grilli1997': {'breakingHeightOverDepth': record['reference']['breakingHeightOverDepth'],
                          'breakingDepthOverDepth': record['reference']['breakingDepthOverDepth']}}
    if 'densityDiffusion' in record:
        out['densityDiffusion'] = record['densityDiffusion']
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), target)
    json.dump(out, open(path, 'w'), indent = 1)
    print(path)


def main(root: str) -> None:
    write(root, 'grilliBreaking_64',
          'BreakingWave(depth = 1, waveHeight = 0.3, beachSlope = 1/35, particlesPerDepth = 64, scheme = deltaSph)',
          'grilli_breaking_64.json')
    write(root, 'grilliBreaking_128_antuono',
          "BreakingWave(depth = 1, waveHeight = 0.3, beachSlope = 1/35, particlesPerDepth = 128, scheme = deltaSph, "
          "densityDiffusion = 'antuono')", 'grilli_breaking_128_antuono.json')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
