'''

The host-side scenario code's outputs, frozen: every configuration's start, the beds and height
fields, the boundaries' layers and clamp, jittered lattices, a split and a merge, and the breaking
read-out with two trackers' whole histories, against the SHA-256 digests in
tests/golden/host_freeze.json (written by tests/golden/make_host_freeze.py). Every golden file of
the breaking tests and every oracle comparison starts from these states, so they must stay bit for
bit what they were when the goldens were made.

'''

import json
import os
import sys

import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


@pytest.mark.timeout(300)
def testTheHostOutputsAreTheFrozenOnes():
    sys.path.insert(0, GOLDEN)
    from make_host_freeze import everything
    frozen = json.load(open(os.path.join(GOLDEN, 'host_freeze.json')))
    record = everything()
    assert sorted(record) == sorted(frozen)
    moved = [key for key in frozen if frozen[key] != record[key]]
    assert not moved, f'outputs moved: {moved}'
