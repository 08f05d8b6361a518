'''

Builds libsph_hip.so, the engine's HIP library, in-tree for gfx950 with hipcc.

Packed float32 instructions are switched off (-packed-fp32-ops), and float32 division and square
root are correctly rounded where the source writes them: the neighbor predicate's root inside its
rounding band, so the float32 predicate is the one the float32 restatement evaluates on the host.
The pair sums call the hardware root and reciprocal by name.

'''

from __future__ import annotations

import os
import shutil
import subprocess

here = os.path.dirname(os.path.abspath(__file__))
root = os.path.dirname(here)
library = os.path.join(here, 'libsph_hip.so')
sources = [os.path.join(here, 'csrc', 'sph_hip.hip')]
headers = [os.path.join(here, 'csrc', 'sph_device.hpp'), os.path.join(root, 'include', 'sph_hip.h')]

flags = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared', '-Wall',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-fhip-fp32-correctly-rounded-divide-sqrt', '-ffp-contract=fast',
         '-I', os.path.join(root, 'include')]


def hipcc() -> str:
    '''The hipcc compiler driver.'''
    found = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(found):
        raise RuntimeError('hipcc was not found; the HIP library cannot be built.')
    return found


def build(verbose: bool = False, output: str | None = None, defines: tuple = ()) -> str:
    '''Compile the HIP library (to `output`, with extra -D `defines`); returns its path.'''
    target = output or library
    temporary = target + '.tmp'
    command = [hipcc()] + flags + ['-D' + define for define in defines] + ['-o', temporary] + sources
    result = subprocess.run(command, capture_output = True, text = True)
    if result.returncode != 0:
        raise RuntimeError('hipcc failed:\n' + result.stdout + result.stderr)
    if verbose and result.stderr:
        print(result.stderr)
    os.replace(temporary, target)
    return target


if __name__ == '__main__':
    print(build(verbose = True))
