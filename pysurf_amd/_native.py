'''

ctypes binding of libsph_hip.so (include/sph_hip.h).

The library must be built (`__graft_entry__.build()` or `python -m pysurf_amd.build`); loading
fails loudly when it is missing or when no HIP device is present, since the engine has no
other path to run on.

'''

from __future__ import annotations

import ctypes
import os

from pysurf_amd.errors import EngineError

here = os.path.dirname(os.path.abspath(__file__))
# SPH_HIP_LIBRARY names another build of the same library (a variant compiled with other launch constants)
libraryPath = os.environ.get('SPH_HIP_LIBRARY') or os.path.join(here, 'libsph_hip.so')

SPH_OK, SPH_ERR_VALUE, SPH_ERR_HIP = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)


class Boundary(ctypes.Structure):
    '''sph_boundary'''
    _fields_ = [('kind', ctypes.c_int32), ('clearance', ctypes.c_double), ('lower', ctypes.c_double * 3),
                ('upper', ctypes.c_double * 3), ('openTop', ctypes.c_int32), ('nBed', ctypes.c_int32),
                ('bedX', ctypes.c_void_p), ('bedZ', ctypes.c_void_p), ('sideWalls', ctypes.c_int32),
                ('width', ctypes.c_double), ('offshoreWall', ctypes.c_int32), ('wallX', ctypes.c_double),
                ('pistonSpeed', ctypes.c_double), ('endLimit', ctypes.c_double), ('fieldNx', ctypes.c_int32),
                ('fieldNy', ctypes.c_int32), ('table', ctypes.c_void_p), ('originX', ctypes.c_double),
                ('originY', ctypes.c_double), ('stepX', ctypes.c_double), ('stepY', ctypes.c_double),
                ('lowY', ctypes.c_double), ('highY', ctypes.c_double), ('lowX', ctypes.c_double),
                ('highX', ctypes.c_double)]


class Run(ctypes.Structure):
    '''sph_run'''
    _fields_ = [('first', ctypes.c_int64), ('count', ctypes.c_int64), ('kernel', ctypes.c_int32),
                ('grid', ctypes.c_int32 * 3), ('cell', ctypes.c_double), ('hmax', ctypes.c_double),
                ('hmin', ctypes.c_double), ('period', ctypes.c_double), ('rho0', ctypes.c_double),
                ('c0', ctypes.c_double), ('alpha', ctypes.c_double), ('boundaryFactor', ctypes.c_double),
                ('gravity', ctypes.c_double * 3), ('diffusion', ctypes.c_int32), ('delta', ctypes.c_double),
                ('extrapolatedWalls', ctypes.c_int32), ('negativePressure', ctypes.c_int32),
                ('correction', ctypes.c_int32), ('limiter', ctypes.c_int32), ('densityFloor', ctypes.c_double),
                ('cfl', ctypes.c_double), ('maxTimeStep', ctypes.c_double), ('boundary', Boundary),
                ('filterInterval', ctypes.c_int32)]


class State(ctypes.Structure):
    '''sph_state'''
    _fields_ = [('position', ctypes.c_void_p), ('velocity', ctypes.c_void_p), ('density', ctypes.c_void_p),
                ('pressure', ctypes.c_void_p), ('acceleration', ctypes.c_void_p), ('mass', ctypes.c_void_p),
                ('length', ctypes.c_void_p), ('fluid', ctypes.c_void_p), ('stride', ctypes.c_int64)]


class ProfileSpec(ctypes.Structure):
    '''sph_profile_spec'''
    _fields_ = [('x0', ctypes.c_double), ('width', ctypes.c_double), ('columns', ctypes.c_int32),
                ('offset', ctypes.c_int32), ('stripCentre', ctypes.c_double), ('stripHalfWidth', ctypes.c_double)]


class FaceSpec(ctypes.Structure):
    '''sph_face_spec'''
    _fields_ = [('windowStart', ctypes.c_double), ('windowEnd', ctypes.c_double), ('base', ctypes.c_double),
                ('spacing', ctypes.c_double), ('columns', ctypes.c_int32), ('bands', ctypes.c_int32),
                ('offset', ctypes.c_int32), ('active', ctypes.c_int32), ('wetHeight', ctypes.c_double),
                ('bedX', ctypes.c_void_p), ('bedZ', ctypes.c_void_p), ('bedPoints', ctypes.c_int32),
                ('stripCentre', ctypes.c_double), ('stripHalfWidth', ctypes.c_double)]


class FractionSpec(ctypes.Structure):
    '''sph_fraction_spec'''
    _fields_ = [('x0', ctypes.c_double), ('y0', ctypes.c_double), ('stepX', ctypes.c_double), ('stepY', ctypes.c_double),
                ('binWidth', ctypes.c_double), ('nx', ctypes.c_int32), ('ny', ctypes.c_int32), ('reach', ctypes.c_int32),
                ('pad', ctypes.c_int32), ('offset', ctypes.c_int64), ('h', ctypes.c_double), ('area', ctypes.c_double),
                ('areaPerMass', ctypes.c_double), ('lowX', ctypes.c_double), ('lowZ', ctypes.c_double),
                ('highX', ctypes.c_double), ('highZ', ctypes.c_double), ('stripCentre', ctypes.c_double),
                ('stripHalfWidth', ctypes.c_double)]


class SplitSpec(ctypes.Structure):
    '''sph_split_spec'''
    _fields_ = [('active', ctypes.c_int32), ('pad', ctypes.c_int32), ('lowX', ctypes.c_double), ('highX', ctypes.c_double),
                ('minMass', ctypes.c_double)]


class MergeSpec(ctypes.Structure):
    '''sph_merge_spec'''
    _fields_ = [('active', ctypes.c_int32), ('pad', ctypes.c_int32), ('lowX', ctypes.c_double), ('highX', ctypes.c_double),
                ('referenceMass', ctypes.c_double), ('referenceDensity', ctypes.c_double), ('lengthRatio', ctypes.c_double)]


# Every entry point include/sph_hip.h declares, with its argument types
signatures = {
    'sph_create': [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p],
    'sph_destroy': [ctypes.c_void_p],
    'sph_last_error': [ctypes.c_void_p],
    'sph_configure': [ctypes.c_void_p, ctypes.POINTER(Run), ctypes.c_int32, ctypes.c_int64],
    'sph_bind': [ctypes.c_void_p, ctypes.POINTER(State)],
    'sph_build_cells': [ctypes.c_void_p],
    'sph_wall_pressure': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'sph_velocity_gradients': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'sph_correction_matrices': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'sph_pair_accelerations': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'sph_continuity_rate': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'sph_shepard_density': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'sph_density_gradients': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'sph_set_closure': [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double],
    'sph_shifting_displacement': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int32,
                                  ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'sph_closure_coefficients': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'sph_xsph_velocity': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p],
    'sph_neighbor_counts': [ctypes.c_void_p, ctypes.c_void_p],
    'sph_neighbor_rows': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'sph_enforce': [ctypes.c_void_p],
    'sph_set_clock': [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_int64,
                      ctypes.c_int64],
    'sph_get_clock': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'sph_run_failures': [ctypes.c_void_p, ctypes.c_void_p],
    'sph_advance': [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32],
    'sph_rates': [ctypes.c_void_p, ctypes.c_void_p],
    'sph_resident_rates': [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p],
    'sph_set_energy_budget': [ctypes.c_void_p, ctypes.c_int32],
    'sph_energy_sums': [ctypes.c_void_p, ctypes.c_void_p],
    'sph_integrate': [ctypes.c_void_p, ctypes.c_double],
    'sph_synchronize': [ctypes.c_void_p],
    'sph_global_sums': [ctypes.c_void_p, ctypes.c_void_p],
    'sph_sort_state': [ctypes.c_void_p, ctypes.c_void_p],
    'sph_surface_profile': [ctypes.c_void_p, ctypes.POINTER(ProfileSpec), ctypes.c_void_p],
    'sph_face_cells': [ctypes.c_void_p, ctypes.POINTER(FaceSpec), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'sph_fluid_fraction': [ctypes.c_void_p, ctypes.POINTER(FractionSpec), ctypes.c_void_p],
    'sph_split_count': [ctypes.c_void_p, ctypes.POINTER(SplitSpec), ctypes.c_void_p, ctypes.c_void_p],
    'sph_split_apply': [ctypes.c_void_p, ctypes.POINTER(State), ctypes.c_double, ctypes.c_double],
    'sph_merge': [ctypes.c_void_p, ctypes.POINTER(MergeSpec), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    'sph_set_launch': [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32],
    'sph_set_near_cells': [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32],
    'sph_profile': [ctypes.c_void_p, ctypes.c_int32],
    'sph_kernel_times': [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32],
    'sph_state_sums': [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Run), ctypes.c_int32, ctypes.POINTER(State),
                       ctypes.c_void_p, ctypes.c_void_p],
    'sph_static_error': [],
}

# The columns of sph_global_sums, per run
globalSumNames = ['mass', 'kinetic', 'potential', 'elastic', 'momentumX', 'momentumY', 'momentumZ']

# The columns of sph_energy_sums, per run: the energy budget's device terms (sphEnergy.py:92-98)
energyTermNames = ['pressureFluid', 'viscosityFluid', 'pressureWall', 'viscosityWall', 'gravityWork', 'continuityFluid',
                   'continuityWall', 'correction', 'diffusion', 'floor', 'drift', 'clamp', 'shifting', 'filter']

# The failure codes of sph_run_failures, and what each means
failureNames = {0: '', 1: 'non-finite', 2: 'time step collapsed'}

# Kernel ids of sph_kernel_times
kernelNames = ['clock', 'cellKeys', 'cellScan', 'cellScatter', 'cellRank', 'gatherRecords', 'sweepWallAndGradients',
               'sweepAccelerations', 'kick', 'sweepContinuity', 'buildNeighbors', 'enforce', 'maxima']

_library = None


def load(path: str | None = None):
    '''The loaded HIP library with its signatures set; raises EngineError when it is not built.'''
    global _library
    if _library is not None and path is None:
        return _library
    path = path or libraryPath
    if not os.path.exists(path):
        raise EngineError(f'The HIP library {path} is not built; run __graft_entry__.build() or '
                          'python -m pysurf_amd.build first.')
    library = ctypes.CDLL(path)
    for name, arguments in signatures.items():
        function = getattr(library, name)
        function.argtypes = arguments
        function.restype = ctypes.c_char_p if name in ('sph_last_error', 'sph_static_error') else ctypes.c_int
    if path == libraryPath:
        _library = library
    return library


def check(status: int, context, what: str) -> None:
    '''Raise the error a status maps onto: ValueError for SPH_ERR_VALUE, EngineError otherwise.'''
    if status == SPH_OK:
        return
    message = load().sph_last_error(context)
    text = (message.decode() if message else '') or what
    if status == SPH_ERR_VALUE:
        raise ValueError(text)
    raise EngineError(f'{what}: {text}')
