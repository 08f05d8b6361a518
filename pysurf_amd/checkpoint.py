'''

Checkpoints of a batch: a versioned binary file from which a fresh engine built on the same runs
resumes bit for bit. It takes the place of the reference's pickled solver state
(BreakingWave3D.py:785-858), whose resume the reference holds to bit-exactness
(tests/testBreakingWave3D.py:269-296); nothing in it is executed on loading.

The file, little-endian throughout:

    magic      8 bytes   b'SPHCKPT' and a zero byte
    version    u32       FORMAT_VERSION
    header     u32       the length of the header in bytes
    header     UTF-8 JSON: the batch's dimensions, precision, launch shape, runs (slot ranges and
                         particle counts), clocks, the caller's `extra` and the arrays' names,
                         dtypes and shapes, in file order
    arrays     each one raw, in C order, starting at a multiple of 16 bytes

The arrays are the engine's state in slot order as the device holds it (after a device sort, too),
with the slot-to-particle identity when the batch has been sorted. The header also keeps a digest
of each run's physics and boundary, and a restore into runs of other physics (sound speed,
smoothing length, diffusion, gravity, ...) or over another bed or height field is refused, as the
reference refuses a checkpoint whose settings or bed hash differ (BreakingWave3D.py:578-597,
842-849).

'''

from __future__ import annotations

import hashlib
import json
import os

import numpy as np

MAGIC = b'SPHCKPT\x00'
FORMAT_VERSION = 2
ALIGN = 16
DTYPES = {'float64': np.float64, 'float32': np.float32, 'int32': np.int32, 'int64': np.int64, 'uint8': np.uint8}


def writeCheckpoint(path: str, header: dict, arrays: dict) -> None:
    '''Write `header` (JSON-serialisable) and the named numpy `arrays` to `path`, atomically.'''
    entries = []
    for name, values in arrays.items():
        values = np.ascontiguousarray(values)
        dtype = values.dtype.name
        if dtype not in DTYPES:
            raise ValueError(f'Array {name!r} has dtype {dtype}, which a checkpoint does not store.')
        entries.append({'name': name, 'dtype': dtype, 'shape': list(values.shape)})
    document = json.dumps({**header, 'arrays': entries}, sort_keys = True).encode('utf-8')
    temporary = path + '.partial'
    with open(temporary, 'wb') as f:
        f.write(MAGIC)
        f.write(np.uint32(FORMAT_VERSION).tobytes())
        f.write(np.uint32(len(document)).tobytes())
        f.write(document)
        for values in arrays.values():
            f.write(b'\x00' * (-f.tell() % ALIGN))
            values = np.ascontiguousarray(values)
            f.write(values.astype(values.dtype.newbyteorder('<'), copy = False).tobytes())
    os.replace(temporary, path)


def readCheckpoint(path: str) -> tuple[dict, dict]:
    '''The header and the named arrays of a checkpoint file.'''
    with open(path, 'rb') as f:
        if f.read(len(MAGIC)) != MAGIC:
            raise ValueError(f'{path} is not a checkpoint.')
        version = int(np.frombuffer(f.read(4), dtype = '<u4')[0])
        if version != FORMAT_VERSION:
            raise ValueError(f'{path} is checkpoint format {version}; this build reads format {FORMAT_VERSION}.')
        length = int(np.frombuffer(f.read(4), dtype = '<u4')[0])
        header = json.loads(f.read(length).decode('utf-8'))
        arrays = {}
        for entry in header['arrays']:
            f.read(-f.tell() % ALIGN)
            dtype = np.dtype(DTYPES[entry['dtype']]).newbyteorder('<')
            count = int(np.prod(entry['shape'], dtype = np.int64))
            data = f.read(count * dtype.itemsize)
            if len(data) != count * dtype.itemsize:
                raise ValueError(f'{path} ends inside array {entry["name"]!r}.')
            arrays[entry['name']] = np.frombuffer(data, dtype = dtype).reshape(entry['shape']).astype(dtype.newbyteorder('='))
    return header, arrays


STATE = ('position', 'velocity', 'density', 'pressure', 'acceleration', 'mass', 'length', 'fluid')


def runDigest(case) -> str:

    '''

    SHA-256 of what fixes a run's step apart from its particles: the run's sph_run as the engine is
    given it (kernel, grid, lengths, physics, options, boundary), with its first slot and the
    boundary tables' addresses blanked, followed by the values of those tables (the bed polyline,
    or the height field).

    '''

    import torch

    from pysurf_amd.batch import runOf
    from pysurf_amd.pair_sums import boundaryOf
    boundary, tables = boundaryOf(case.boundary, case.dimensions, torch.device('cpu'))
    run = runOf(case, 0, boundary)
    run.boundary.bedX = run.boundary.bedZ = run.boundary.table = None
    digest = hashlib.sha256(bytes(run))
    for table in tables:
        digest.update(table.numpy().tobytes())
    return digest.hexdigest()


def save(batch, path: str, extra: dict | None = None) -> None:

    '''

    Write `batch`'s state, clocks and launch shape to `path`. `extra` is any JSON-serialisable
    record the caller keeps with it (a tracker's history, say). Synchronizes.

    '''

    engine = batch.engine
    engine.synchronize()
    count = batch.count
    arrays = {}
    for name in STATE:
        tensor = engine.tensors[name]
        arrays[name] = (tensor[:, :count] if tensor.dim() == 2 else tensor[:count]).cpu().numpy()
    if batch.identity is not None:
        arrays['identity'] = batch.identity[:count].cpu().numpy()
    clock = batch.clocks()
    runs = len(batch.cases)
    header = {
        'dimensions': batch.dimensions, 'precision': batch.precision, 'count': count,
        'runs': [{'first': batch.offsets[r], 'count': batch.offsets[r + 1] - batch.offsets[r]} for r in range(runs)],
        'digests': [runDigest(case) for case in batch.cases],
        'launch': {key: int(batch.launch.get(key, -1))
                   for key in ('lanes', 'block', 'listCapacity', 'nearCells', 'nearCellsBuild')},
        'clock': {'time': [float(t).hex() for t in clock['time']], 'steps': [int(s) for s in clock['steps']],
                  'until': [float(u).hex() for u in batch.until], 'maxSteps': [int(m) for m in batch.maxSteps]},
        'extra': extra if extra is not None else {},
    }
    writeCheckpoint(path, header, arrays)


def restore(batch, path: str) -> dict:

    '''

    Load the checkpoint at `path` into `batch`, a BatchEngine built on the same runs as the one
    saved, and return the caller's `extra`. Raises ValueError when the batch does not match the
    checkpoint's runs, dimensions or precision, or when a run's physics or boundary differ from
    the saved run's.

    '''

    import torch
    header, arrays = readCheckpoint(path)
    runs = [{'first': batch.offsets[r], 'count': batch.offsets[r + 1] - batch.offsets[r]} for r in range(len(batch.cases))]
    if header['dimensions'] != batch.dimensions or header['precision'] != batch.precision or header['runs'] != runs:
        raise ValueError('The checkpoint was written by a batch of other runs, dimensions or precision.')
    digests = [runDigest(case) for case in batch.cases]
    differ = [r for r, (saved, now) in enumerate(zip(header['digests'], digests)) if saved != now]
    if differ:
        raise ValueError(f'The checkpoint was written for other physics or another bed: runs {differ} differ from '
                         'the saved runs (settings or bed hash).')
    launch = dict(header['launch'])
    launch.setdefault('nearCells', -1)
    launch.setdefault('nearCellsBuild', -1)
    batch.engine.setLaunch(launch['block'], launch['lanes'], launch['listCapacity'])
    batch.engine.setNearCells(launch['nearCells'], launch['nearCellsBuild'])
    batch.launch = launch
    engine, count = batch.engine, header['count']
    with torch.cuda.device(engine.device), torch.cuda.stream(engine.stream):
        for name in STATE:
            tensor = engine.tensors[name]
            source = torch.from_numpy(np.ascontiguousarray(arrays[name])).to(tensor.device)
            if source.dtype != tensor.dtype:
                raise ValueError(f'The checkpoint holds {name} as {source.dtype}, the engine as {tensor.dtype}.')
            (tensor[:, :count] if tensor.dim() == 2 else tensor[:count]).copy_(source)
        if 'identity' in arrays:
            identity = torch.arange(max(count, 1), dtype = torch.int32, device = engine.device)
            identity[:count].copy_(torch.from_numpy(np.ascontiguousarray(arrays['identity'])).to(engine.device))
            batch.identity = identity
        else:
            batch.identity = None
    engine.synchronize()
    clock = header['clock']
    times = [float.fromhex(t) for t in clock['time']]
    until = [float.fromhex(u) for u in clock['until']]
    for r in range(len(runs)):
        engine.setClock(r, times[r], until[r], clock['steps'][r], clock['maxSteps'][r])
    batch.until, batch.maxSteps = until, list(clock['maxSteps'])
    return header['extra']
