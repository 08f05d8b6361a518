'''

Builds libsph_hip.so, the engine's HIP library, in-tree for gfx950 with hipcc.

Packed float32 instructions are switched off (-packed-fp32-ops), and float32 division and square
root are correctly rounded where the source writes them: the neighbor predicate's root inside its
rounding band, so the float32 predicate is the one the float32 restatement evaluates on the host.
The pair sums call the hardware root and reciprocal by name.

Each source compiles to its own object, all at once, and the objects link into the one library.

'''

from __future__ import annotations

import os
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

here = os.path.dirname(os.path.abspath(__file__))
root = os.path.dirname(here)
library = os.path.join(here, 'libsph_hip.so')
sources = [os.path.join(here, 'csrc', name) for name in ('sph_hip.hip', 'sph_static.hip')]
headers = [os.path.join(here, 'csrc', name) for name in ('sph_device.hpp', 'sph_sums.hpp')] + \
          [os.path.join(root, 'include', 'sph_hip.h')]

flags = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-Wall',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-fhip-fp32-correctly-rounded-divide-sqrt', '-ffp-contract=fast',
         '-I', os.path.join(root, 'include')]


def hipcc() -> str:
    '''The hipcc compiler driver.'''
    found = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(found):
        raise RuntimeError('hipcc was not found; the HIP library cannot be built.')
    return found


def _run(command: list) -> str:
    result = subprocess.run(command, capture_output = True, text = True)
    if result.returncode != 0:
        raise RuntimeError('hipcc failed:\n' + ' '.join(command) + '\n' + result.stdout + result.stderr)
    return result.stderr


def build(verbose: bool = False, output: str | None = None, defines: tuple = ()) -> str:
    '''Compile the HIP library (to `output`, with extra -D `defines`); returns its path.'''
    target = output or library
    temporary = target + '.tmp'
    extra = ['-D' + define for define in defines]
    with tempfile.TemporaryDirectory(prefix = 'sph_build_') as folder:
        objects = [os.path.join(folder, os.path.basename(source) + '.o') for source in sources]
        with ThreadPoolExecutor(max_workers = len(sources)) as pool:
            logs = list(pool.map(lambda pair: _run([hipcc()] + flags + extra + ['-c', '-o', pair[1], pair[0]]),
                                 zip(sources, objects)))
        logs.append(_run([hipcc(), '--offload-arch=gfx950', '-shared', '-fPIC', '-o', temporary] + objects))
    if verbose:
        for log in logs:
            if log:
                print(log)
    os.replace(temporary, target)
    return target


if __name__ == '__main__':
    print(build(verbose = True))
