'''

The checkpoint file format on the host: arrays of every stored dtype come back bit for bit with
their shapes, the header comes back as written, and a file of another format, another version or
cut short is refused with ValueError rather than read.

'''

import numpy as np
import pytest

from pysurf_amd import checkpoint


def arrays():
    rng = np.random.default_rng(4)
    return {
        'position': rng.normal(size = (3, 17)),
        'pressure': rng.normal(size = 17).astype(np.float32),
        'fluid': (rng.uniform(size = 17) < 0.8).astype(np.uint8),
        'identity': rng.permutation(17).astype(np.int32),
        'steps': np.array([3, 2**40], dtype = np.int64),
        'empty': np.zeros((2, 0)),
        'special': np.array([np.nan, np.inf, -0.0, 5e-324], dtype = np.float64),
    }


def testArraysAndHeaderComeBackBitForBit(tmp_path):
    path = str(tmp_path / 'state.ckpt')
    header = {'dimensions': 2, 'clock': {'time': [0.1.hex()]}, 'extra': {'history': [1.5, 2.5]}}
    written = arrays()
    checkpoint.writeCheckpoint(path, header, written)
    readHeader, read = checkpoint.readCheckpoint(path)
    assert {k: v for k, v in readHeader.items() if k != 'arrays'} == header
    assert list(read) == list(written)
    for name, values in written.items():
        assert read[name].dtype == values.dtype and read[name].shape == values.shape, name
        assert read[name].tobytes() == np.ascontiguousarray(values).tobytes(), name


def testOtherFilesAreRefused(tmp_path):
    path = str(tmp_path / 'state.ckpt')
    checkpoint.writeCheckpoint(path, {}, arrays())
    data = open(path, 'rb').read()
    other = tmp_path / 'other'
    other.write_bytes(b'PK\x03\x04' + data[4:])
    with pytest.raises(ValueError, match = 'not a checkpoint'):
        checkpoint.readCheckpoint(str(other))
    newer = bytearray(data)
    newer[8:12] = np.uint32(checkpoint.FORMAT_VERSION + 1).tobytes()
    other.write_bytes(bytes(newer))
    with pytest.raises(ValueError, match = 'format'):
        checkpoint.readCheckpoint(str(other))
    other.write_bytes(data[:-8])
    with pytest.raises(ValueError, match = 'ends inside'):
        checkpoint.readCheckpoint(str(other))


def testUnstoredDtypesAreRefused(tmp_path):
    with pytest.raises(ValueError, match = 'dtype'):
        checkpoint.writeCheckpoint(str(tmp_path / 'x'), {}, {'c': np.zeros(2, dtype = np.complex128)})


def testRunDigestsTellPhysicsAndBedsApart():
    # The digest a checkpoint keeps of each run changes with the bed (the same particle count over a
    # steeper slope) and with the physics (another sound speed), and repeats for the same run
    import dataclasses

    from pysurf_amd.scenarios import solitaryOnSlope
    case = solitaryOnSlope(particlesPerDepth = 6, seed = 1)
    again = solitaryOnSlope(particlesPerDepth = 6, seed = 1)
    assert checkpoint.runDigest(case) == checkpoint.runDigest(again)
    otherBed = dataclasses.replace(case, boundary = solitaryOnSlope(particlesPerDepth = 6, seed = 1, slope = 1.0 / 12.0).boundary)
    assert otherBed.count == case.count
    assert checkpoint.runDigest(otherBed) != checkpoint.runDigest(case)
    louder = dataclasses.replace(case, soundSpeed = case.soundSpeed * 1.01)
    assert checkpoint.runDigest(louder) != checkpoint.runDigest(case)
