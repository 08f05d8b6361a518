'''

The C ABI without a GPU: the built library loads and exports every entry point include/sph_hip.h
declares, the ctypes structures match the header's layout, and the Python binding covers them all.
No compute call runs here.

'''

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'sph_hip.h')


def declared() -> list:
    text = open(HEADER).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char\*)\s+(sph_\w+)\s*\(', text, flags = re.M)))


def testTheLibraryExportsEveryDeclaredEntryPoint():
    from pysurf_amd import _native
    if not os.path.exists(_native.libraryPath):
        pytest.fail('libsph_hip.so is not built; run __graft_entry__.build()')
    library = ctypes.CDLL(_native.libraryPath)
    names = declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(library, name), name
    assert set(names) == set(_native.signatures)


def testStructuresMatchTheHeaderLayout():
    from pysurf_amd import _native
    # sizeof and offsetof as gcc lays out include/sph_hip.h on x86-64
    assert ctypes.sizeof(_native.Boundary) == 216
    assert ctypes.sizeof(_native.Run) == 400
    assert _native.Run.filterInterval.offset == 392
    assert ctypes.sizeof(_native.State) == 72
    assert _native.Run.boundary.offset == 176
    assert ctypes.sizeof(_native.ProfileSpec) == 40
    assert ctypes.sizeof(_native.FaceSpec) == 96
    assert _native.FaceSpec.bedPoints.offset == 72
    assert _native.FaceSpec.stripCentre.offset == 80
    assert ctypes.sizeof(_native.FractionSpec) == 136
    assert _native.FractionSpec.offset.offset == 56
    assert _native.FractionSpec.h.offset == 64
    assert _native.FractionSpec.stripCentre.offset == 120
    assert ctypes.sizeof(_native.SplitSpec) == 32
    assert _native.SplitSpec.minMass.offset == 24
    assert ctypes.sizeof(_native.MergeSpec) == 48
    assert _native.MergeSpec.lengthRatio.offset == 40


def testBadArgumentsFailWithoutADevice():
    from pysurf_amd import _native
    library = _native.load()
    handle = ctypes.c_void_p()
    assert library.sph_create(ctypes.byref(handle), 4, 0, None) == _native.SPH_ERR_VALUE
    assert library.sph_create(ctypes.byref(handle), 2, 7, None) == _native.SPH_ERR_VALUE
    assert library.sph_destroy(None) == _native.SPH_OK
    assert library.sph_set_near_cells(None, 0, -1) == _native.SPH_ERR_VALUE
