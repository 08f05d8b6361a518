'''

Writes tests/golden/grilli_breaking_64.json: the breaking point the reference recorded for Grilli's
1:35 case at 64 particles per depth on its GPU backend, copied from its results file
validation/gpuSph/results/grilliBreaking_64.json (data, read as JSON; no reference code runs), with
Grilli, Svendsen and Subramanya's (1997) values it is judged against.

    python tests/golden/make_grilli_record.py [reference root]

'''

import json
import os
import sys

FIELDS = ('particlesPerDepth', 'bedFitted', 'faceArc', 'fluidParticles', 'breakingTime', 'breakingHeightOverDepth',
          'breakingDepthOverDepth', 'breakingFaceDepthOverDepth', 'breakingTroughOverDepth', 'touchdownTime',
          'touchedDown')


def main(root: str) -> None:
    source = os.path.join(root, 'validation', 'gpuSph', 'results', 'grilliBreaking_64.json')
    record = json.load(open(source))
    summary = record['summary']
    out = {'source': 'validation/gpuSph/results/grilliBreaking_64.json (the reference GPU backend run of '
                     'validation/gpuSph/grilliBreaking.py 64)',
           'case': 'BreakingWave(depth = 1, waveHeight = 0.3, beachSlope = 1/35, particlesPerDepth = 64, scheme = deltaSph)',
           'recorded': {name: summary[name] for name in FIELDS},
           'grilli1997': {'breakingHeightOverDepth': record['reference']['breakingHeightOverDepth'],
                          'breakingDepthOverDepth': record['reference']['breakingDepthOverDepth']}}
    target = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'grilli_breaking_64.json')
    json.dump(out, open(target, 'w'), indent = 1)
    print(target)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
