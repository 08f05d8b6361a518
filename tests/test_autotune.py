'''

The autotuner's candidate shapes and cache key, on the CPU: a batch stand-in carries what the tuner
reads of a BatchEngine before it times anything (the dimension and the runs' smoothing lengths and
spacings).

'''

from types import SimpleNamespace

from pysurf_amd import autotune


def standIn(dimensions: int, ratios: list, count: int = 100_000):
    cases = [SimpleNamespace(smoothingLength = ratio * 0.01, spacing = 0.01) for ratio in ratios]
    return SimpleNamespace(dimensions = dimensions, cases = cases, count = count, precision = 'float32')


def testShortLengthsKeepTheDefaultList():
    # at h = 1.3 dx a 2D particle has about 21 neighbors and a 3D one about 74: the default lists hold them
    assert autotune.candidatesFor(standIn(2, [1.3])) == autotune.CANDIDATES[2]
    assert autotune.candidatesFor(standIn(3, [1.3])) == autotune.CANDIDATES[3]


def testLongLengthsTryLongerLists():
    # at h = 3 dx a 2D particle has about 113 neighbors: lists of 1.5 and 2.5 times that join the shapes
    shapes = autotune.candidatesFor(standIn(2, [1.3, 3.0]))
    capacities = {shape[2] for shape in shapes}
    assert capacities == {-1, 192, 288}
    assert all(len(shape) == 3 or shape[2] == -1 for shape in shapes)
    assert all(len(autotune.full(shape)) == 5 for shape in shapes)
    assert all(capacity % 32 == 0 for capacity in capacities if capacity > 0)


def testTheCacheTellsLengthRatiosApart():
    # two batches of one dimension, precision and size but other smoothing lengths over spacings
    # are tuned apart
    short, long = standIn(2, [1.3]), standIn(2, [3.0])
    keys = [autotune.cacheKey(b.dimensions, b.precision, b.count, autotune.lengthRatios(b)) for b in (short, long)]
    assert keys[0] != keys[1]
    assert autotune.lengthRatios(standIn(2, [1.3, 1.3, 3.0])) == (1.3, 3.0)
