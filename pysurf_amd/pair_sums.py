'''

The MI355X counterpart of pySurf's `sphGpu.PairSums` (src/pysurf/sphGpu.py:1492-2323): the same
method surface over the HIP engine (include/sph_hip.h), with device state held in PyTorch-ROCm
tensors and every sum, step and enforcement run by hand-written HIP kernels.

    host steps      `update`, then `pairAccelerations`, `continuityRate`, `wallPressure`,
                    `velocityGradients`, `correctionMatrices`, `densityGradients`, `xsphVelocity`,
                    `shepardDensity`, `closureCoefficients`, `shiftingDisplacement`:
                    host float64 arrays in, host float64 arrays out, shaped as the reference
                    returns them
    device steps    `setPhysics`, `push`, `setBoundary`, then `residentRates` and
                    `residentIntegrate(dt)` each step, or `advance`, which keeps the time step on
                    the device and replays the steps as a HIP graph; `pull` brings the state back

Pair arithmetic is float32 over float64 state by default, each float64 position carried as a
float32 point and remainder as the reference does (sphGpu.py:28-38); `pairPrecision = 'float64'`
runs the plain float64 kernels of the same source.

The options the delta-SPH scheme leaves off are built as host-step sums: XSPH, the Shepard filter,
the renormalized 'antuono' diffusion (`densityGradients`, and `continuityRate` under
setDiffusion('antuono')), the delta-LES-SPH closure (`setClosure`, `closureCoefficients`, and the
closure's terms in `pairAccelerations` and `continuityRate`) and particle shifting (`setShifting`,
`shiftingDisplacement`). The device steps carry the Shepard filter (`residentRates(True)`, or
`setFilterInterval` for `advance`), the 'antuono' diffusion and the closure as the reference's do
(sphGpu.py:2237-2315), and the energy budget's per-particle ledger (`setEnergyBudget`,
`energySums`, sphGpu.py:1780-1825). XSPH and shifting stay host steps: setPhysics and the device
steps raise ValueError when asked for either.

'''

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from pysurf_amd import _native
from pysurf_amd.errors import EngineError

taitExponent = 7.0
diffusionModes = {'none': 0, 'molteni': 1, 'hydrostatic': 2, 'antuono': 3}
maxGridCells = 2**25


def cubicSplineNormalization(smoothingLength: float, dimensions: int) -> float:
    '''The cubic spline's sigma: 10 / (7 pi h^2) in 2D, 1 / (pi h^3) in 3D (sphGpu.py:1424-1429).'''
    h = smoothingLength
    return 10.0 / (7.0 * math.pi * h * h) if dimensions == 2 else 1.0 / (math.pi * h * h * h)


def wendlandNormalization(smoothingLength: float, dimensions: int) -> float:
    '''The Wendland C2's sigma: 7 / (4 pi h^2) in 2D, 21 / (16 pi h^3) in 3D (sphGpu.py:1431-1436).'''
    h = smoothingLength
    return 7.0 / (4.0 * math.pi * h * h) if dimensions == 2 else 21.0 / (16.0 * math.pi * h * h * h)


def kernelNormalization(kernelType: str, smoothingLength: float, dimensions: int) -> float:
    '''The sigma of 'cubicSpline' or 'wendlandC2' (sphGpu.py:1438-1444).'''
    if kernelType == 'wendlandC2':
        return wendlandNormalization(smoothingLength, dimensions)
    return cubicSplineNormalization(smoothingLength, dimensions)


def gridDimensions(domainMin, domainMax, cellSize: float, dimensions: int) -> tuple:
    '''Cells along x, y and z covering the domain, one support across, capped at 2^25 (sphGpu.py:1466-1478).'''
    extent = np.zeros(3)
    extent[:dimensions] = np.asarray(domainMax, dtype = float) - np.asarray(domainMin, dtype = float)
    cells = np.maximum(np.ceil(extent / cellSize).astype(int) + 3, 1)
    if dimensions == 2:
        cells[2] = 1
    while int(np.prod(cells)) > maxGridCells:
        cells = np.maximum((cells * 0.8).astype(int), 1)
    return tuple(int(count) for count in cells)


def deviceOf(device) -> torch.device:
    '''The HIP device of a Warp-style device name; the engine has no CPU path.'''
    if device is None:
        device = 'cuda:0'
    device = torch.device(device)
    if device.type != 'cuda':
        raise ValueError(f'The HIP engine runs on a GPU, not on {device}.')
    if not torch.cuda.is_available():
        raise EngineError('No HIP device is present; the engine has no CPU path.')
    return device


class Engine:

    '''

    A context of the HIP library over one or more runs, with the runs' state in torch tensors.
    `PairSums` holds one; the batch engine (pysurf_amd.batch) holds one over many runs.

    '''

    def __init__(self, dimensions: int, precision: str, device = None) -> None:
        if dimensions not in (2, 3):
            raise ValueError(f'Dimensions must be 2 or 3, not {dimensions}.')
        if precision not in ('float32', 'float64'):
            raise ValueError(f"Pair precision '{precision}' is not 'float32' or 'float64'.")
        self.dimensions = dimensions
        self.precision = precision
        self.device = deviceOf(device)
        self.real = torch.float32 if precision == 'float32' else torch.float64
        self.library = _native.load()
        with torch.cuda.device(self.device):
            self.stream = torch.cuda.Stream(device = self.device)
            handle = ctypes.c_void_p()
            status = self.library.sph_create(ctypes.byref(handle), dimensions, 0 if precision == 'float32' else 1,
                                             ctypes.c_void_p(self.stream.cuda_stream))
        if status != _native.SPH_OK:
            raise EngineError('sph_create failed: no HIP device or a bad argument.')
        self.handle = handle
        self.capacity = 0
        self.count = 0
        self.tensors: dict[str, torch.Tensor] = {}
        self._runs = None
        self._layout = None
        self._keep: list = []

    def __del__(self) -> None:
        handle = getattr(self, 'handle', None)
        if handle is not None and handle.value:
            try:
                self.library.sph_destroy(handle)
            except Exception:
                pass
            self.handle = None

    def check(self, status: int, what: str) -> None:
        _native.check(status, self.handle, what)

    # -- State -- #
    def reserve(self, count: int) -> None:
        '''State tensors for at least `count` slots, growing by 1.25 as the reference's _reserve does (sphGpu.py:1580-1594).'''
        if count <= self.capacity and self.tensors:
            return
        capacity = max(count, int(1.25 * self.capacity), 1)
        old, oldCount = self.tensors, self.count
        fresh = self._allocate(capacity)
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            if old and oldCount:
                for name, tensor in fresh.items():
                    tensor[..., :oldCount].copy_(old[name][..., :oldCount])
        self.tensors = fresh
        self.capacity = capacity
        self._layout = None

    def _allocate(self, capacity: int) -> dict:
        '''Zeroed state and host-step tensors of `capacity` slots.'''
        d, dev, real = self.dimensions, self.device, self.real
        with torch.cuda.device(dev), torch.cuda.stream(self.stream):
            return {
                'position': torch.zeros((d, capacity), dtype = torch.float64, device = dev),
                'velocity': torch.zeros((d, capacity), dtype = torch.float64, device = dev),
                'density': torch.zeros(capacity, dtype = torch.float64, device = dev),
                'pressure': torch.zeros(capacity, dtype = real, device = dev),
                'acceleration': torch.zeros((d, capacity), dtype = real, device = dev),
                'mass': torch.zeros(capacity, dtype = real, device = dev),
                'length': torch.zeros(capacity, dtype = real, device = dev),
                'fluid': torch.zeros(capacity, dtype = torch.uint8, device = dev),
                # host-step inputs and outputs
                'velocityIn': torch.zeros((d, capacity), dtype = torch.float64, device = dev),
                'densityIn': torch.zeros(capacity, dtype = torch.float64, device = dev),
                'pressureIn': torch.zeros(capacity, dtype = real, device = dev),
                'vector': torch.zeros((d, capacity), dtype = real, device = dev),
                'scalar': torch.zeros(capacity, dtype = real, device = dev),
                'scalar64': torch.zeros(capacity, dtype = torch.float64, device = dev),
                'matrix': torch.zeros((d * d, capacity), dtype = real, device = dev),
            }

    def stateOf(self, tensors: dict, stride: int) -> _native.State:
        t = tensors
        return _native.State(t['position'].data_ptr(), t['velocity'].data_ptr(), t['density'].data_ptr(),
                             t['pressure'].data_ptr(), t['acceleration'].data_ptr(), t['mass'].data_ptr(),
                             t['length'].data_ptr(), t['fluid'].data_ptr(), stride)

    def split(self, specs: list, referenceDensity: float, lengthRatio: float = 1.0, mask = None) -> np.ndarray:

        '''

        Split the fluid particles `specs` (one _native.SplitSpec per run) and `mask` (a uint8 tensor
        over the slots, or None) select into 2^dims daughters on the device (sph_split_count,
        sph_split_apply; splitParticles, sphRefinement.py:135-186), growing each run in place of
        the layout: run r's own particles first, then its mothers' other daughters. Returns how
        many particles of each run split. The runs' clocks restart at zero; the caller sets them
        again. Synchronizes.

        '''

        runs = len(self._runs)
        array = self._readSpecs(_native.SplitSpec, specs)
        chosen = np.zeros(runs, dtype = np.int64)
        maskPointer = ctypes.c_void_p(mask.data_ptr()) if mask is not None else None
        self.check(self.library.sph_split_count(self.handle, array, maskPointer, chosen.ctypes.data_as(ctypes.c_void_p)),
                   'sph_split_count')
        if not chosen.any():
            return chosen
        grown = chosen * (2**self.dimensions - 1)
        total = self.count + int(grown.sum())
        target = self.tensors if total <= self.capacity else self._allocate(max(total, int(1.25 * self.capacity)))
        stride = self.capacity if target is self.tensors else target['position'].shape[1]
        self.check(self.library.sph_split_apply(self.handle, ctypes.byref(self.stateOf(target, stride)),
                                                float(referenceDensity), float(lengthRatio)), 'sph_split_apply')
        fresh = []
        first = 0
        for r in range(runs):
            run = _native.Run.from_buffer_copy(self._runs[r])
            run.first, run.count = first, int(self._runs[r].count) + int(grown[r])
            if chosen[r] and lengthRatio < 1.0:
                run.hmin = min(run.hmin, run.hmin * lengthRatio)
            first += run.count
            fresh.append(run)
        if target is not self.tensors:
            self.tensors, self.capacity = target, stride
        self.configure(fresh)
        return chosen

    def merge(self, specs: list, mask = None) -> tuple[np.ndarray, np.ndarray]:

        '''

        Merge the split fluid particles `specs` (one _native.MergeSpec per run) and `mask` (a uint8
        tensor over the slots, or None) allow, 2^dims of one level at a time, on the device
        (sph_merge; mergeParticles, sphRefinement.py:212-311), the runs closing up in place.
        Returns each run's groups merged and kinetic energy removed. The runs' clocks restart at
        zero; the caller sets them again. Synchronizes.

        '''

        runs = len(self._runs)
        array = self._readSpecs(_native.MergeSpec, specs)
        groups = np.zeros(runs, dtype = np.int64)
        energy = np.zeros(runs)
        maskPointer = ctypes.c_void_p(mask.data_ptr()) if mask is not None else None
        self.check(self.library.sph_merge(self.handle, array, maskPointer, groups.ctypes.data_as(ctypes.c_void_p),
                                          energy.ctypes.data_as(ctypes.c_void_p)), 'sph_merge')
        if groups.any():
            shrunk = groups * (2**self.dimensions - 1)
            fresh, first = [], 0
            for r in range(runs):
                run = _native.Run.from_buffer_copy(self._runs[r])
                run.first, run.count = first, int(self._runs[r].count) - int(shrunk[r])
                first += run.count
                fresh.append(run)
            self.configure(fresh)
        return groups, energy

    def configure(self, runs: list) -> None:
        '''Hand the runs (a list of _native.Run) to the library and bind the state tensors.'''
        count = sum(int(run.count) for run in runs)
        self.reserve(count)
        self.count = count
        array = (_native.Run * len(runs))(*runs)
        self._runs = array
        self.check(self.library.sph_configure(self.handle, array, len(runs), self.capacity), 'sph_configure')
        state = self.stateOf(self.tensors, self.capacity)
        self.check(self.library.sph_bind(self.handle, ctypes.byref(state)), 'sph_bind')
        # a new layout of the runs starts the ledger afresh
        if getattr(self, 'energyBudget', False):
            self.setEnergyBudget(True)

    def upload(self, name: str, values, count: int | None = None) -> torch.Tensor:
        '''Copy host values (N,), or (N, dim) into the rows of the field, into a state tensor.'''
        count = self.count if count is None else count
        tensor = self.tensors[name]
        values = np.asarray(values)
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            source = torch.from_numpy(np.ascontiguousarray(values.T if values.ndim == 2 else values))
            if tensor.dim() == 2:
                tensor[:values.shape[1] if values.ndim == 2 else 1, :count].copy_(source.to(tensor.dtype), non_blocking = False)
            else:
                tensor[:count].copy_(source.to(tensor.dtype), non_blocking = False)
        return tensor

    def uploadSlice(self, name: str, values, first: int) -> None:
        '''Copy host values (n,) or (n, dim) into slots [first, first + n) of a state tensor.'''
        tensor = self.tensors[name]
        values = np.asarray(values)
        count = values.shape[0]
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            source = torch.from_numpy(np.ascontiguousarray(values.T if values.ndim == 2 else values))
            if tensor.dim() == 2:
                rows = values.shape[1] if values.ndim == 2 else 1
                tensor[:rows, first:first + count].copy_(source.to(tensor.dtype), non_blocking = False)
            else:
                tensor[first:first + count].copy_(source.to(tensor.dtype), non_blocking = False)

    def copySlice(self, name: str, tensor: torch.Tensor, source: int, first: int, count: int) -> None:
        '''Copy slots [source, source + count) of a device tensor laid out as this field's into slots
        [first, first + count) of the field, on the device.'''
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            target = self.tensors[name]
            rows = tensor[..., source:source + count]
            if target.dim() == 2:
                target[:rows.shape[0], first:first + count].copy_(rows)
            else:
                target[first:first + count].copy_(rows)

    def download(self, name: str, rows: int | None = None, count: int | None = None) -> np.ndarray:
        '''A state tensor's first `count` slots as float64, (N, rows) for a field of rows, (N,) otherwise.'''
        count = self.count if count is None else count
        self.synchronize()
        tensor = self.tensors[name]
        if tensor.dim() == 2:
            rows = tensor.shape[0] if rows is None else rows
            return tensor[:rows, :count].to(torch.float64).cpu().numpy().T.copy()
        return tensor[:count].to(torch.float64).cpu().numpy()

    def synchronize(self) -> None:
        self.check(self.library.sph_synchronize(self.handle), 'sph_synchronize')

    def pointer(self, name: str) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.tensors[name].data_ptr())

    # -- Clock -- #
    def setClock(self, run: int, time: float, untilTime: float, steps: int, maxSteps: int) -> None:
        self.check(self.library.sph_set_clock(self.handle, run, float(time), float(untilTime), int(steps),
                                              int(maxSteps)), 'sph_set_clock')

    def clock(self, runs: int) -> dict:
        time, dt = np.zeros(runs), np.zeros(runs)
        steps = np.zeros(runs, dtype = np.int64)
        maxima = np.zeros(2 * runs)
        self.check(self.library.sph_get_clock(self.handle, time.ctypes.data_as(ctypes.c_void_p),
                                              dt.ctypes.data_as(ctypes.c_void_p),
                                              steps.ctypes.data_as(ctypes.c_void_p),
                                              maxima.ctypes.data_as(ctypes.c_void_p)), 'sph_get_clock')
        failed = np.zeros(runs, dtype = np.int32)
        self.check(self.library.sph_run_failures(self.handle, failed.ctypes.data_as(ctypes.c_void_p)), 'sph_run_failures')
        return {'time': time, 'dt': dt, 'steps': steps, 'maxima': maxima.reshape(runs, 2), 'failed': failed}

    def advance(self, steps: int, graph: bool = True) -> None:
        self.check(self.library.sph_advance(self.handle, int(steps), 1 if graph else 0), 'sph_advance')

    # -- Options of the device step -- #
    def setClosure(self, enabled: bool, soundSpeed: float = 1.0, referenceDensity: float = 1000.0,
                   kinematicViscosity: float = 1e-6) -> None:
        '''The delta-LES-SPH closure of every run (sph_set_closure).'''
        self.check(self.library.sph_set_closure(self.handle, 1 if enabled else 0, float(soundSpeed),
                                                float(referenceDensity), float(kinematicViscosity)), 'sph_set_closure')

    def setEnergyBudget(self, enabled: bool) -> None:
        '''Switch the per-particle energy ledger on (cleared to zero) or off (sph_set_energy_budget).'''
        self.energyBudget = bool(enabled)
        self.check(self.library.sph_set_energy_budget(self.handle, 1 if enabled else 0), 'sph_set_energy_budget')

    def energySums(self, runs: int) -> np.ndarray:
        '''Each run's ledger totals (sph_energy_sums), shape (runs, 14) in _native.energyTermNames order.'''
        sums = np.zeros((runs, len(_native.energyTermNames)))
        self.check(self.library.sph_energy_sums(self.handle, sums.ctypes.data_as(ctypes.c_void_p)), 'sph_energy_sums')
        return sums

    def sortState(self, identity: torch.Tensor | None = None) -> None:
        '''Permute the state into cell order within each run, carrying `identity` (int32 per slot) along (sph_sort_state).'''
        pointer = ctypes.c_void_p(identity.data_ptr()) if identity is not None else None
        self.check(self.library.sph_sort_state(self.handle, pointer), 'sph_sort_state')

    def globalSums(self, runs: int) -> np.ndarray:
        '''Each run's float64 sums over its fluid (sph_global_sums), shape (runs, 7) in _native.globalSumNames order.'''
        sums = np.zeros((runs, len(_native.globalSumNames)))
        self.check(self.library.sph_global_sums(self.handle, sums.ctypes.data_as(ctypes.c_void_p)), 'sph_global_sums')
        return sums

    # -- Breaking read-out -- #
    def _readSpecs(self, kind, specs: list):
        if self._runs is None or len(specs) != len(self._runs):
            raise ValueError(f'The read-out needs one specification per run ({len(specs)} given).')
        return (kind * len(specs))(*specs)

    def surfaceProfile(self, specs: list) -> np.ndarray:
        '''

        The highest tracked particle of each column (sph_surface_profile): `specs` holds one
        _native.ProfileSpec per run; returns the tops, -inf where a column is empty. Synchronizes.

        '''
        array = self._readSpecs(_native.ProfileSpec, specs)
        total = max([spec.offset + spec.columns for spec in specs if spec.columns > 0], default = 0)
        tops = np.empty(max(total, 1))
        self.check(self.library.sph_surface_profile(self.handle, array, tops.ctypes.data_as(ctypes.c_void_p)),
                   'sph_surface_profile')
        return tops[:total]

    def faceCells(self, specs: list) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        '''

        The face windows' marked cells and the most landward particles near the bed
        (sph_face_cells): `specs` holds one _native.FaceSpec per run. Returns the occupancy words,
        each run's overflow flag and its landward points, (runs, 8, 2) by descending x, NaN past the
        last. Synchronizes.

        '''
        array = self._readSpecs(_native.FaceSpec, specs)
        words = max([spec.offset + spec.bands * -(-spec.columns // 32) for spec in specs if spec.active], default = 0)
        occupancy = np.zeros(max(words, 1), dtype = np.uint32)
        overflow = np.zeros(len(specs), dtype = np.int32)
        landward = np.zeros((len(specs), 8, 2))
        self.check(self.library.sph_face_cells(self.handle, array, occupancy.ctypes.data_as(ctypes.c_void_p),
                                               overflow.ctypes.data_as(ctypes.c_void_p),
                                               landward.ctypes.data_as(ctypes.c_void_p)), 'sph_face_cells')
        return occupancy[:words], overflow, landward

    def fluidFraction(self, specs: list) -> np.ndarray:
        '''

        The runs' smoothed fluid fractions (sph_fluid_fraction): `specs` holds one
        _native.FractionSpec per run; returns every run's nodes, each at its offset. Synchronizes.

        '''
        array = self._readSpecs(_native.FractionSpec, specs)
        total = max([spec.offset + spec.nx * spec.ny for spec in specs if spec.nx > 0], default = 0)
        fraction = np.zeros(max(total, 1))
        self.check(self.library.sph_fluid_fraction(self.handle, array, fraction.ctypes.data_as(ctypes.c_void_p)),
                   'sph_fluid_fraction')
        return fraction[:total]

    def setLaunch(self, blockSize: int = 256, lanes: int = 0, listCapacity: int = -1) -> None:
        '''

        Threads per workgroup and lanes per particle of the pair sweeps (0 lanes: 8 in 2D, 16 in 3D),
        and the neighbor list's entries per particle (0: every sweep walks the cells; -1: 64 in 2D,
        160 in 3D).

        '''
        self.check(self.library.sph_set_launch(self.handle, int(blockSize), int(lanes), int(listCapacity)),
                   'sph_set_launch')

    def setNearCells(self, bytes: int = -1, buildBytes: int = -1) -> None:
        '''The cap of a sweep workgroup's LDS copy of the records around its particles (-1: 32 KiB), and
        the same for the device step's list build (-1: `bytes`).'''
        self.check(self.library.sph_set_near_cells(self.handle, int(bytes), int(buildBytes)), 'sph_set_near_cells')


def boundaryOf(boundary, dimensions: int, device: torch.device, pistonSpeed: float = 0.0) -> tuple:
    '''

    The sph_boundary of a ContainerBoundary, BedBoundary or HeightFieldBoundary, as
    PairSums.setBoundary lays it out (sphGpu.py:2121-2171); returns it with the device tables it
    points into, which the caller keeps alive.

    '''

    from pysurf_amd import boundaries

    BedBoundary, ContainerBoundary = boundaries.BedBoundary, boundaries.ContainerBoundary
    HeightFieldBoundary = getattr(boundaries, 'HeightFieldBoundary', type(None))
    b = _native.Boundary()
    keep = []
    if boundary is None:
        b.kind = 0
    elif isinstance(boundary, ContainerBoundary):
        b.kind = 1
        lower, upper = np.zeros(3), np.zeros(3)
        lower[:dimensions] = boundary.containerMin
        upper[:dimensions] = boundary.containerMax
        b.lower[:] = list(lower)
        b.upper[:] = list(upper)
        b.openTop = 1 if boundary.openTop else 0
    elif isinstance(boundary, BedBoundary):
        bed = boundary.bathymetry
        b.kind = 2
        b.clearance = 0.25 * boundary.spacing
        bedX = torch.tensor(bed.x, dtype = torch.float64, device = device)
        bedZ = torch.tensor(bed.z, dtype = torch.float64, device = device)
        keep += [bedX, bedZ]
        b.nBed = len(bed.x)
        b.bedX, b.bedZ = bedX.data_ptr(), bedZ.data_ptr()
        b.sideWalls = 1 if (boundary.width is not None and not boundary.periodic) else 0
        b.width = float(boundary.width or 0.0)
        b.offshoreWall = 1 if boundary.wallTop is not None else 0
        b.wallX = float(boundary.wallX)
        b.pistonSpeed = float(pistonSpeed)
        b.endLimit = bed.endX - b.clearance if boundary.endWallTop is not None else bed.endX
    elif isinstance(boundary, HeightFieldBoundary):
        field = boundary.field
        b.kind = 3
        b.clearance = 0.25 * boundary.spacing
        table = torch.tensor(np.ascontiguousarray(field.z), dtype = torch.float64, device = device)
        keep.append(table)
        b.fieldNy, b.fieldNx = field.z.shape
        b.table = table.data_ptr()
        startX, endX, startY, endY = field.bounds
        b.originX, b.originY = float(field.x[0]), float(field.y[0])
        b.stepX, b.stepY = float(field.spacingX), float(field.spacingY)
        b.lowY, b.highY = startY + b.clearance, endY - b.clearance
        b.offshoreWall = 1 if boundary.offshoreWall else 0
        b.lowX = startX + b.clearance
        b.highX = endX - b.clearance if boundary.endWall else endX
    else:
        raise ValueError(f'The device cannot enforce a {type(boundary).__name__}; step from the host.')
    return b, keep


class PairSums:

    '''

    The SPH pair sums and the time step of one particle system on an MI355X
    (sphGpu.PairSums, sphGpu.py:1492-1570).

    Parameters:
    -----------
    dimensions : int
        2 or 3
    smoothingLength : float
        h [m]; the support is 2h
    domainMin, domainMax : np.ndarray
        Corners of the region the particles occupy [m], which sizes the cell grid
    period : float
        Length of a periodic y axis in 3D [m]; 0 for none
    separationPrecision : str
        'float64', separations from each float64 position's float32 point and remainder, the
        only form the engine builds
    device : str | None
        'cuda:N'; None for cuda:0
    kernelType : str
        'cubicSpline' or 'wendlandC2'
    pairPrecision : str
        'float32' pair sums over float64 state, or 'float64' throughout (the plain kernel)

    '''

    def __init__(self, dimensions: int, smoothingLength: float, domainMin: np.ndarray, domainMax: np.ndarray,
                 period: float = 0.0, separationPrecision: str = 'float64', device: str | None = None,
                 kernelType: str = 'cubicSpline', pairPrecision: str = 'float32') -> None:

        if separationPrecision not in ('float64', 'float32'):
            raise ValueError(f"Separation precision '{separationPrecision}' is not 'float64' or 'float32'.")
        if separationPrecision == 'float32':
            raise ValueError("The engine takes separations from the float64 positions only ('float64').")
        if kernelType not in ('cubicSpline', 'wendlandC2'):
            raise ValueError(f"Kernel '{kernelType}' is not 'cubicSpline' or 'wendlandC2'.")
        if period > 0.0 and dimensions != 3:
            raise ValueError('A periodic axis needs three dimensions: it is y.')
        self.dimensions = dimensions
        self.smoothingLength = float(smoothingLength)
        self.kernelType = kernelType
        self.kind = 1 if kernelType == 'wendlandC2' else 0
        self.sigma = kernelNormalization(kernelType, smoothingLength, dimensions)
        self.sigmaUnit = kernelNormalization(kernelType, 1.0, dimensions)
        self.maxLength = self.smoothingLength
        self.minLength = self.smoothingLength
        self.uniformLengths = True
        self.period = float(period)
        self.precise = 1
        self.pairPrecision = pairPrecision
        self.gridShape = gridDimensions(domainMin, domainMax, 2.0 * smoothingLength, dimensions)
        self.count = 0
        self.physics: dict = {}
        self.diffusion = {'mode': 0, 'factor': 0.0, 'delta': 0.1, 'referenceDensity': 1000.0,
                          'gravityOverSound': 0.0, 'soundSpeed': 1.0}
        self.extrapolateWalls = False
        self.negativePressure = False
        self.correctContinuity = False
        self.limitViscosity = False
        self.closure = {'enabled': False, 'soundSpeed': 1.0, 'referenceDensity': 1000.0, 'kinematicViscosity': 0.0}
        self.shifting = {'enabled': False, 'mach': 0.0, 'spacingKernel': 0.0, 'soundSpeed': 1.0, 'local': False}
        self.trackEnergy = False
        self.filterInterval = 0
        self.engine = Engine(dimensions, pairPrecision, device)
        self.device = self.engine.device
        self._boundary = _native.Boundary()
        self._boundaryKeep: list = []
        self._configured = None

    # -- Configuration -- #
    def _run(self) -> _native.Run:
        '''This system's sph_run from its settings.'''
        run = _native.Run()
        run.first, run.count = 0, self.count
        run.kernel = self.kind
        run.grid[:] = list(self.gridShape)
        run.cell = 2.0 * self.smoothingLength
        run.hmax, run.hmin = self.maxLength, self.minLength
        run.period = self.period
        physics = self.physics
        run.rho0 = physics.get('referenceDensity', self.diffusion['referenceDensity'])
        run.c0 = physics.get('soundSpeed', self.diffusion['soundSpeed'])
        run.alpha = physics.get('alpha', 0.0)
        run.boundaryFactor = physics.get('boundaryViscosityFactor', 0.0)
        g = np.zeros(3)
        g[:self.dimensions] = physics.get('gravity64', np.zeros(self.dimensions))
        run.gravity[:] = list(g)
        run.diffusion = self.diffusion['mode']
        run.delta = self.diffusion['delta']
        run.extrapolatedWalls = int(self.extrapolateWalls)
        run.negativePressure = int(self.negativePressure)
        run.correction = int(self.correctContinuity)
        run.limiter = int(self.limitViscosity)
        run.densityFloor = physics.get('densityFloor', 0.0)
        run.cfl = physics.get('cfl', 0.25)
        run.maxTimeStep = physics.get('maxTimeStep', 1.0)
        run.boundary = self._boundary
        run.filterInterval = self.filterInterval
        return run

    def _configure(self) -> None:
        '''Hand the current settings to the engine (the cells are rebuilt by the next update or step).'''
        self.engine.configure([self._run()])
        self._configured = True

    def setDiffusion(self, mode: str, delta: float, soundSpeed: float, referenceDensity: float, gravity: float) -> None:
        '''The delta-SPH term the continuity sum carries (sphGpu.py:1879-1904).'''
        if mode not in diffusionModes:
            raise ValueError(f"Density diffusion '{mode}' is not 'none', 'molteni', 'hydrostatic' or 'antuono'.")
        self.diffusion = {'mode': diffusionModes[mode], 'factor': 2.0 * delta * soundSpeed, 'delta': float(delta),
                          'referenceDensity': float(referenceDensity), 'gravityOverSound': gravity / soundSpeed**2,
                          'soundSpeed': float(soundSpeed)}
        self._configured = None

    def setWallPressure(self, extrapolated: bool) -> None:
        '''Whether boundary particles take their pressure from the fluid (sphGpu.py:1906-1914).'''
        self.extrapolateWalls = bool(extrapolated)
        self._configured = None

    def setNegativePressure(self, enabled: bool) -> None:
        '''Whether the fluid may carry tension (sphGpu.py:1916-1920).'''
        self.negativePressure = bool(enabled)
        self._configured = None

    def setContinuityCorrection(self, enabled: bool) -> None:
        '''Whether the continuity sum takes the kernel gradient correction (sphGpu.py:1696-1702).'''
        self.correctContinuity = bool(enabled)
        self._configured = None

    def setViscosityLimiter(self, enabled: bool) -> None:
        '''Whether the artificial viscosity takes the limited velocity difference (sphGpu.py:1734-1740).'''
        self.limitViscosity = bool(enabled)
        self._configured = None

    def setShifting(self, enabled: bool, mach: float = 0.1, spacing: float = 0.0, soundSpeed: float = 1.0,
                    local: bool = False) -> None:

        '''

        Particle shifting (sphGpu.py:1924-1955) for `shiftingDisplacement`: Ma in the shift's size
        CFL Ma (2h)^2, the spacing dx for the anti-clumping weight's W(dx), c0 for the step's Courant
        number, or with `local` each particle's own size A h |v_i| dt. The device steps refuse it;
        the delta-SPH scheme runs without it.

        '''

        if enabled and not self.uniformLengths:
            raise ValueError('Particle shifting needs one smoothing length for every particle.')
        if enabled and not spacing > 0.0:
            raise ValueError('Particle shifting needs the particle spacing.')
        spacingKernel = 0.0
        if enabled:
            q = spacing / self.smoothingLength
            if self.kernelType == 'wendlandC2':
                spacingKernel = self.sigma * max(1.0 - 0.5 * q, 0.0)**4 * (2.0 * q + 1.0)
            else:
                spacingKernel = self.sigma * ((1.0 - 1.5 * q**2 + 0.75 * q**3) if q < 1.0
                                              else 0.25 * max(2.0 - q, 0.0)**3)
        self.shifting = {'enabled': bool(enabled), 'mach': float(mach), 'spacingKernel': spacingKernel,
                         'soundSpeed': float(soundSpeed), 'local': bool(local)}

    def shiftingDisplacement(self, densities: np.ndarray, timeStep: float,
                             velocities: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:

        '''

        This step's shift of every particle, zero on the boundary and the free surface [m], and the
        density [kg/m^3] and velocity [m/s] changes that go with it (sphGpu.py:1966-1979).

        '''

        shifting, h = self.shifting, self.smoothingLength
        if not shifting['enabled']:
            raise ValueError('Particle shifting is off; see setShifting.')
        if shifting['local']:
            size = 2.0 * h * timeStep
        else:
            size = (shifting['soundSpeed'] * timeStep / h) * shifting['mach'] * (2.0 * h)**2
        self._sync()
        engine = self.engine
        engine.upload('velocityIn', velocities)
        engine.upload('densityIn', densities)
        # shifts into 'vector', density changes into 'scalar', velocity changes into the first rows of 'matrix'
        self.check(engine.library.sph_shifting_displacement(
            engine.handle, engine.pointer('velocityIn'), engine.pointer('densityIn'), float(size),
            1 if shifting['local'] else 0, float(shifting['spacingKernel']), engine.pointer('vector'),
            engine.pointer('scalar'), engine.pointer('matrix')), 'sph_shifting_displacement')
        dim = self.dimensions
        velocity = engine.download('matrix')[:, :dim]
        return engine.download('vector'), engine.download('scalar'), velocity

    def setClosure(self, enabled: bool, soundSpeed: float = 1.0, referenceDensity: float = 1000.0,
                   kinematicViscosity: float = 1e-6) -> None:

        '''

        Whether the accelerations and diffusive term carry the delta-LES-SPH closure
        (sphGpu.py:1754-1778): each fluid pair's viscosity adds K nu / (c0 h) plus the harmonic mean
        of the eddy alphas, and the diffusive term takes the harmonic mean of the eddy deltas, in
        the host steps and in the device steps alike.

        '''

        self.closure = {'enabled': bool(enabled), 'soundSpeed': float(soundSpeed),
                        'referenceDensity': float(referenceDensity), 'kinematicViscosity': float(kinematicViscosity)}
        self.engine.setClosure(enabled, soundSpeed, referenceDensity, kinematicViscosity)

    def closureCoefficients(self, velocities: np.ndarray, densities: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        '''Each particle's eddy alpha and delta at the current positions (sphGpu.py:1841-1853).'''
        if not self.closure['enabled']:
            raise ValueError('The closure is off; see setClosure.')
        self._sync()
        engine = self.engine
        engine.upload('velocityIn', velocities)
        engine.upload('densityIn', densities)
        deltas = torch.empty(max(self.count, 1), dtype = engine.tensors['scalar'].dtype, device = self.device)
        self.check(engine.library.sph_closure_coefficients(engine.handle, engine.pointer('velocityIn'),
                                                           engine.pointer('densityIn'), engine.pointer('scalar'),
                                                           ctypes.c_void_p(deltas.data_ptr())),
                   'sph_closure_coefficients')
        engine.synchronize()
        return engine.download('scalar'), deltas[:self.count].double().cpu().numpy()

    def setEnergyBudget(self, enabled: bool) -> None:

        '''

        Whether the device steps keep the energy budget (sphGpu.py:1780-1793): each fluid particle's
        running float64 total of every term in `_native.energyTermNames` (sphEnergy.deviceTerms),
        from `push` to `energySums`.

        '''

        self.trackEnergy = bool(enabled)
        if self.count:
            self.engine.setEnergyBudget(self.trackEnergy)

    def setFilterInterval(self, interval: int) -> None:
        '''The Shepard filter every `interval` steps of `advance` (densityFilterInterval, sphSolver.py:1979-1981); 0 never.'''
        if interval < 0:
            raise ValueError('The filter interval cannot be negative.')
        self.filterInterval = int(interval)
        self._configured = None

    # -- Host steps -- #
    def update(self, positions: np.ndarray, masses: np.ndarray, isFluid: np.ndarray,
               lengths: np.ndarray | None = None) -> None:

        '''

        Copy the positions, masses, fluid flags and smoothing lengths to the device and rebuild the
        cell list (sphGpu.py:1998-2042).

        Raises:
        -------
        ValueError : If a smoothing length exceeds the grid's.

        '''

        count = len(positions)
        if lengths is None:
            lengths = np.full(count, self.smoothingLength)
        lengths = np.asarray(lengths, dtype = float)
        if count and lengths.max() > self.maxLength * (1.0 + 1e-9):
            raise ValueError(f'A smoothing length of {lengths.max()} m exceeds the {self.maxLength} m the cell grid '
                             f'is built for.')
        self.count = count
        if count:
            self.minLength = float(lengths.min())
            self.uniformLengths = bool(np.all(lengths == lengths[0]))
        self._configure()
        if count == 0:
            return
        engine = self.engine
        engine.upload('position', np.asarray(positions, dtype = float))
        engine.upload('mass', np.asarray(masses, dtype = float))
        engine.upload('length', lengths)
        engine.upload('fluid', np.asarray(isFluid, dtype = bool).astype(np.uint8))
        self.check(engine.library.sph_build_cells(engine.handle), 'sph_build_cells')

    def check(self, status: int, what: str) -> None:
        self.engine.check(status, what)

    def _sync(self) -> None:
        """Push changed settings to the engine: a descriptor upload, the layout and cells unchanged."""
        if self._configured is None and self.count:
            self.check(self.engine.library.sph_configure(self.engine.handle, ctypes.byref(self._run()), 1,
                                                         self.engine.capacity), 'sph_configure')
            self._configured = True

    def _gravityPhysics(self, soundSpeed, alpha, boundaryFactor, gravity, referenceDensity = None) -> None:
        physics = dict(self.physics)
        physics.update(soundSpeed = float(soundSpeed), alpha = float(alpha),
                       boundaryViscosityFactor = float(boundaryFactor),
                       gravity64 = np.asarray(gravity, dtype = float))
        if referenceDensity is not None:
            physics['referenceDensity'] = float(referenceDensity)
        self.physics = physics
        self._configured = None
        self._sync()

    def pairAccelerations(self, velocities: np.ndarray, densities: np.ndarray, pressures: np.ndarray,
                          soundSpeed: float, alpha: float, boundaryViscosityFactor: float,
                          gravity: np.ndarray) -> np.ndarray:
        '''Pressure gradient, artificial viscosity and gravity on the fluid, zero on the boundary (sphGpu.py:2044-2059).'''
        self._gravityPhysics(soundSpeed, alpha, boundaryViscosityFactor, gravity)
        engine = self.engine
        engine.upload('velocityIn', velocities)
        engine.upload('densityIn', densities)
        engine.upload('pressureIn', pressures)
        self.check(engine.library.sph_pair_accelerations(engine.handle, engine.pointer('velocityIn'),
                                                         engine.pointer('densityIn'), engine.pointer('pressureIn'),
                                                         engine.pointer('vector')), 'sph_pair_accelerations')
        return engine.download('vector')

    def continuityRate(self, velocities: np.ndarray, densities: np.ndarray | None = None) -> np.ndarray:
        '''d rho / dt for every particle, with the diffusive term setDiffusion set (sphGpu.py:2061-2078).'''
        if densities is None:
            if self.diffusion['mode'] != 0 or self.correctContinuity:
                raise ValueError('The diffusive term and the correction need the densities.')
            densities = np.full(self.count, self.diffusion['referenceDensity'])
        if self.diffusion['mode'] != 0 and not self.physics:
            self._gravityPhysics(self.diffusion['soundSpeed'], 0.0, 0.0,
                                 self._diffusionGravity(), self.diffusion['referenceDensity'])
        self._sync()
        engine = self.engine
        engine.upload('velocityIn', velocities)
        engine.upload('densityIn', densities)
        self.check(engine.library.sph_continuity_rate(engine.handle, engine.pointer('velocityIn'),
                                                      engine.pointer('densityIn'), engine.pointer('scalar')),
                   'sph_continuity_rate')
        return engine.download('scalar')

    def densityGradients(self, densities: np.ndarray) -> np.ndarray:
        '''Each particle's renormalized density gradient, shape (N, dim) (sphGpu.py:1868-1880).'''
        if not self.renormalizedDiffusion:
            raise ValueError("The diffusive term is not 'antuono'; see setDiffusion.")
        self._sync()
        engine = self.engine
        engine.upload('densityIn', densities)
        self.check(engine.library.sph_density_gradients(engine.handle, engine.pointer('densityIn'),
                                                        engine.pointer('vector')), 'sph_density_gradients')
        return engine.download('vector')

    def xsphVelocity(self, velocities: np.ndarray, densities: np.ndarray, epsilon: float) -> np.ndarray:
        '''The XSPH transport velocity of the fluid; boundary rows unchanged (sphGpu.py:2080-2088).'''
        self._sync()
        engine = self.engine
        engine.upload('velocityIn', velocities)
        engine.upload('densityIn', densities)
        self.check(engine.library.sph_xsph_velocity(engine.handle, engine.pointer('velocityIn'),
                                                    engine.pointer('densityIn'), float(epsilon),
                                                    engine.pointer('vector')), 'sph_xsph_velocity')
        return engine.download('vector')

    def shepardDensity(self, densities: np.ndarray) -> np.ndarray:
        '''The Shepard-filtered density of every particle (sphGpu.py:2104-2112).'''
        self._sync()
        engine = self.engine
        engine.upload('densityIn', densities)
        self.check(engine.library.sph_shepard_density(engine.handle, engine.pointer('densityIn'),
                                                      engine.pointer('scalar')), 'sph_shepard_density')
        return engine.download('scalar')

    def _diffusionGravity(self) -> np.ndarray:
        '''Gravity along the last axis, from setDiffusion's g / c0^2.'''
        g = np.zeros(self.dimensions)
        g[-1] = -self.diffusion['gravityOverSound'] * self.diffusion['soundSpeed']**2
        return g

    def wallPressure(self, densities: np.ndarray, pressures: np.ndarray, gravity: np.ndarray, referenceDensity: float,
                     soundSpeed: float) -> np.ndarray:
        '''Every pressure with the boundary's taken from the fluid (sphGpu.py:2090-2102).'''
        physics = self.physics
        self._gravityPhysics(soundSpeed, physics.get('alpha', 0.0), physics.get('boundaryViscosityFactor', 0.0),
                             gravity, referenceDensity)
        engine = self.engine
        engine.upload('densityIn', densities)
        engine.upload('pressureIn', pressures)
        self.check(engine.library.sph_wall_pressure(engine.handle, engine.pointer('densityIn'),
                                                    engine.pointer('pressureIn'), engine.pointer('scalar'),
                                                    engine.pointer('scalar64')), 'sph_wall_pressure')
        self.wallDensities = engine.download('scalar64')
        return engine.download('scalar')

    def velocityGradients(self, velocities: np.ndarray, densities: np.ndarray) -> np.ndarray:
        '''G for every particle, shape (N, dim, dim) (sphGpu.py:1827-1839).'''
        if not self.needsVelocityGradients:
            raise ValueError('Neither the viscosity limiter nor the closure is on; see setViscosityLimiter, setClosure.')
        self._sync()
        engine = self.engine
        engine.upload('velocityIn', velocities)
        engine.upload('densityIn', densities)
        self.check(engine.library.sph_velocity_gradients(engine.handle, engine.pointer('velocityIn'),
                                                         engine.pointer('densityIn'), engine.pointer('matrix')),
                   'sph_velocity_gradients')
        dim = self.dimensions
        return engine.download('matrix').reshape(self.count, dim, dim)

    @property
    def renormalizedDiffusion(self) -> bool:
        '''Whether the diffusive term subtracts the renormalized density gradients (sphGpu.py:1654-1657).'''
        return self.diffusion['mode'] == diffusionModes['antuono']

    @property
    def needsCorrections(self) -> bool:
        '''Whether a step builds L: for the continuity correction or the density gradients (sphGpu.py:1659-1662).'''
        return self.correctContinuity or self.renormalizedDiffusion

    @property
    def needsVelocityGradients(self) -> bool:
        '''Whether a step builds G: for the viscosity limiter or the closure (sphGpu.py:1664-1667).'''
        return self.limitViscosity or self.closure['enabled']

    def correctionMatrices(self, densities: np.ndarray) -> np.ndarray:
        '''L for every particle, shape (N, dim, dim) (sphGpu.py:1855-1866).'''
        if not self.needsCorrections:
            raise ValueError('Neither the continuity correction nor the renormalized diffusion is on.')
        self._sync()
        engine = self.engine
        engine.upload('densityIn', densities)
        self.check(engine.library.sph_correction_matrices(engine.handle, engine.pointer('densityIn'),
                                                          engine.pointer('matrix')), 'sph_correction_matrices')
        dim = self.dimensions
        return engine.download('matrix').reshape(self.count, dim, dim)

    def neighborLists(self) -> tuple[np.ndarray, np.ndarray]:
        '''Each particle's neighbors {j != i : r_ij < h_i + h_j}, as CSR (offsets, rows), each row ascending.'''
        self._sync()
        engine, n = self.engine, self.count
        counts = torch.zeros(max(n, 1), dtype = torch.int32, device = self.device)
        self.check(engine.library.sph_neighbor_counts(engine.handle, ctypes.c_void_p(counts.data_ptr())),
                   'sph_neighbor_counts')
        counts = counts[:n].cpu().numpy().astype(np.int64)
        offsets = np.zeros(n + 1, dtype = np.int64)
        np.cumsum(counts, out = offsets[1:])
        deviceOffsets = torch.from_numpy(offsets).to(self.device)
        rows = torch.zeros(max(int(offsets[-1]), 1), dtype = torch.int32, device = self.device)
        self.check(engine.library.sph_neighbor_rows(engine.handle, ctypes.c_void_p(deviceOffsets.data_ptr()),
                                                    ctypes.c_void_p(rows.data_ptr())), 'sph_neighbor_rows')
        rows = rows[:int(offsets[-1])].cpu().numpy()
        owner = np.repeat(np.arange(n), counts)
        order = np.lexsort((rows, owner))
        return offsets, rows[order]

    # -- Device steps -- #
    def canEnforce(self, boundary) -> bool:
        '''Whether the device can enforce this boundary (sphGpu.py:2115-2119).'''
        from pysurf_amd import boundaries
        kinds = tuple(getattr(boundaries, name) for name in ('ContainerBoundary', 'BedBoundary', 'HeightFieldBoundary')
                      if hasattr(boundaries, name))
        return boundary is None or isinstance(boundary, kinds)

    def setBoundary(self, boundary, particles) -> None:
        '''The boundary the device steps enforce, a piston held at its speed (sphGpu.py:2121-2171).'''
        pistonSpeed = 0.0
        mask = getattr(boundary, '_pistonMask', None)
        if mask is not None and mask.any():
            pistonSpeed = float(particles.velocities[mask, 0][0])
        self._boundary, self._boundaryKeep = boundaryOf(boundary, self.dimensions, self.device, pistonSpeed)
        self._configured = None
        self._sync()

    def setPhysics(self, referenceDensity: float, soundSpeed: float, alpha: float, boundaryViscosityFactor: float,
                   gravity: np.ndarray, xsphEpsilon: float, densityFloor: float, cflNumber: float = 0.25,
                   maxTimeStep: float = 1.0) -> None:
        '''The constants of the device step (sphGpu.py:2173-2205); XSPH is off under delta-SPH and left out of it.'''
        if xsphEpsilon > 0.0:
            raise ValueError('The device step does not blend in XSPH; the delta-SPH scheme moves the fluid with its '
                             'velocity. xsphVelocity gives the blend as a host step.')
        self.physics = {'referenceDensity': float(referenceDensity), 'soundSpeed': float(soundSpeed),
                        'stiffness': float(referenceDensity * soundSpeed**2 / taitExponent), 'alpha': float(alpha),
                        'boundaryViscosityFactor': float(boundaryViscosityFactor),
                        'gravity64': np.asarray(gravity, dtype = float), 'xsphEpsilon': 0.0,
                        'densityFloor': float(densityFloor), 'cfl': float(cflNumber),
                        'maxTimeStep': float(maxTimeStep)}
        self._configured = None

    def push(self, particles) -> None:
        '''Copy a particle system's state to the device (sphGpu.py:2207-2221).'''
        self.update(particles.positions, particles.masses, particles.isFluid, particles.lengths)
        if self.count == 0:
            return
        self.engine.upload('velocity', np.asarray(particles.velocities, dtype = float))
        self.engine.upload('density', np.asarray(particles.densities, dtype = float))
        # the ledger counts from the state pushed (sphGpu.py:2218-2219)
        if self.trackEnergy:
            self.engine.setEnergyBudget(True)
        self.engine.synchronize()

    def pull(self, particles) -> None:
        '''Copy the device state back into a particle system, in place (sphGpu.py:2223-2235).'''
        if self.count == 0:
            return
        engine, dim = self.engine, self.dimensions
        particles.positions[:] = engine.download('position')[:, :dim]
        particles.velocities[:] = engine.download('velocity')[:, :dim]
        particles.densities[:] = engine.download('density')
        particles.pressures[:] = engine.download('pressure')
        particles.accelerations[:] = engine.download('acceleration')[:, :dim]

    def energySums(self) -> np.ndarray:

        '''

        Each term's energy over the device steps since `push` [J], summed over the particles, in
        `_native.energyTermNames` order (sphGpu.py:1805-1825).

        Raises:
        -------
        ValueError : While the budget is off.

        '''

        if not self.trackEnergy:
            raise ValueError('The energy budget is off; see setEnergyBudget.')
        if self.count == 0:
            return np.zeros(len(_native.energyTermNames))
        return self.engine.energySums(1)[0]

    def globalSums(self) -> dict:

        '''

        Float64 sums over the fluid of the device-resident state, reduced on the device in a fixed
        order: 'mass', 'kinetic' (sum m |v|^2 / 2, ParticleSystem.kineticEnergy), 'potential'
        (sum m g z, ParticleSystem.potentialEnergy with g the downward gravity), 'elastic' (sum m e(rho),
        sphEnergy.elasticEnergy) and 'momentum' (sum m v).

        '''

        sums = self.engine.globalSums(1)[0]
        values = dict(zip(_native.globalSumNames[:4], map(float, sums[:4])))
        values['momentum'] = sums[4:4 + self.dimensions].copy()
        return values

    def residentRates(self, filterDensity: bool = False) -> tuple[float, float]:

        '''

        The first half of a device step: the cells, the Shepard filter of the fluid density when
        `filterDensity`, then the pressure and the accelerations; returns the largest fluid speed
        and acceleration (sphGpu.py:2237-2273).

        '''

        self._refuseShifting()
        self._sync()
        maxima = np.zeros(2)
        self.check(self.engine.library.sph_resident_rates(self.engine.handle, 1 if filterDensity else 0,
                                                          maxima.ctypes.data_as(ctypes.c_void_p)),
                   'sph_resident_rates')
        return float(maxima[0]), float(maxima[1])

    def _refuseShifting(self) -> None:
        if self.shifting['enabled']:
            raise ValueError('The device step does not shift the particles; shiftingDisplacement gives the shift '
                             'as a host step.')

    def residentIntegrate(self, timeStep: float) -> None:
        '''The second half of a device step with the host's time step (sphGpu.py:2275-2315).'''
        self._refuseShifting()
        self.check(self.engine.library.sph_integrate(self.engine.handle, float(timeStep)), 'sph_integrate')

    def enforceBoundary(self) -> None:
        '''Enforce the boundary set by setBoundary on the device state (sphGpu.py:2317-2323).'''
        self._sync()
        self.check(self.engine.library.sph_enforce(self.engine.handle), 'sph_enforce')

    def advance(self, time: float, untilTime: float, steps: int, maxSteps: int, chunk: int = 64,
                graph: bool = True) -> dict:

        '''

        Device steps until the run's time reaches `untilTime` or it has taken `maxSteps` steps, the
        time step reduced on the device, in graphs of `chunk` steps; returns the clock.

        '''

        if self.count == 0:
            return {'time': float(time), 'steps': int(steps), 'dt': 0.0}
        self._refuseShifting()
        self._sync()
        engine = self.engine
        engine.setClock(0, time, untilTime, steps, maxSteps)
        while True:
            engine.advance(chunk, graph)
            clock = engine.clock(1)
            if not (clock['time'][0] < untilTime and clock['steps'][0] < maxSteps):
                return {'time': float(clock['time'][0]), 'steps': int(clock['steps'][0]),
                        'dt': float(clock['dt'][0])}
