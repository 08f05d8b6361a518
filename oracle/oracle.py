'''

ctypes wrapper of the float64 CPU restatement of the delta-SPH step (oracle/sph_oracle.cpp).

TEST INFRASTRUCTURE, NOT PRODUCT CODE: only tests/, `__graft_entry__.smoke()` and bench.py's
cpu_baseline leg may import this module, as the checker or the timed CPU baseline.

Arrays go in as (N, dim) float64 and are padded to (N, 3) here; results come back as the
reference's host-step surface returns them (sphGpu.py:1827-2113): (N, dim), (N,) or (N, dim, dim).

'''

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_here = os.path.dirname(os.path.abspath(__file__))
_libraryPath = os.path.join(_here, 'liboracle.so')


class OracleParams(ctypes.Structure):
    _fields_ = [('dims', ctypes.c_int32), ('kernel', ctypes.c_int32), ('grid', ctypes.c_int32 * 3),
                ('cell', ctypes.c_double), ('hmax', ctypes.c_double), ('hmin', ctypes.c_double),
                ('period', ctypes.c_double), ('rho0', ctypes.c_double), ('c0', ctypes.c_double),
                ('alpha', ctypes.c_double), ('boundaryFactor', ctypes.c_double), ('gravity', ctypes.c_double * 3),
                ('diffusion', ctypes.c_int32), ('delta', ctypes.c_double), ('extrapolatedWalls', ctypes.c_int32),
                ('negativePressure', ctypes.c_int32), ('correction', ctypes.c_int32), ('limiter', ctypes.c_int32),
                ('densityFloor', ctypes.c_double), ('cfl', ctypes.c_double), ('maxTimeStep', ctypes.c_double)]


_dp = ctypes.POINTER(ctypes.c_double)


class OracleBoundary(ctypes.Structure):
    _fields_ = [('kind', ctypes.c_int32), ('clearance', ctypes.c_double), ('lower', ctypes.c_double * 3),
                ('upper', ctypes.c_double * 3), ('openTop', ctypes.c_int32), ('nBed', ctypes.c_int32),
                ('bedX', _dp), ('bedZ', _dp), ('sideWalls', ctypes.c_int32), ('width', ctypes.c_double),
                ('offshoreWall', ctypes.c_int32), ('wallX', ctypes.c_double), ('pistonSpeed', ctypes.c_double),
                ('endLimit', ctypes.c_double), ('fieldNx', ctypes.c_int32), ('fieldNy', ctypes.c_int32),
                ('table', _dp), ('originX', ctypes.c_double), ('originY', ctypes.c_double),
                ('stepX', ctypes.c_double), ('stepY', ctypes.c_double), ('lowY', ctypes.c_double),
                ('highY', ctypes.c_double), ('lowX', ctypes.c_double), ('highX', ctypes.c_double)]


class OracleClosure(ctypes.Structure):
    _fields_ = [('alphas', _dp), ('deltas', _dp), ('soundSpeed', ctypes.c_double),
                ('referenceDensity', ctypes.c_double), ('waterTerm', ctypes.c_double)]


class OracleOptions(ctypes.Structure):
    _fields_ = [('filterInterval', ctypes.c_int32), ('closure', ctypes.c_int32),
                ('closureSoundSpeed', ctypes.c_double), ('closureDensity', ctypes.c_double),
                ('closureWater', ctypes.c_double)]


# The energy budget's device terms, in the ledger's column order (sphEnergy.py deviceTerms).
ledgerTerms = ('pressureFluid', 'viscosityFluid', 'pressureWall', 'viscosityWall', 'gravityWork', 'continuityFluid',
               'continuityWall', 'correction', 'diffusion', 'floor', 'drift', 'clamp', 'shifting', 'filter')


def closureStruct(physics, closure: dict | None):
    '''The OracleClosure of {'alphas', 'deltas', 'soundSpeed', 'referenceDensity', 'kinematicViscosity'}, and
    the arrays it points at; (None, None) without a closure.'''
    if closure is None:
        return None, None
    alphas = np.ascontiguousarray(closure['alphas'], dtype = float)
    deltas = np.ascontiguousarray(closure['deltas'], dtype = float)
    c = OracleClosure()
    c.alphas = alphas.ctypes.data_as(_dp)
    c.deltas = deltas.ctypes.data_as(_dp)
    c.soundSpeed = closure['soundSpeed']
    c.referenceDensity = closure['referenceDensity']
    c.waterTerm = 2.0 * (physics.dims + 2) * closure['kinematicViscosity'] / closure['soundSpeed']
    return c, (alphas, deltas)


def build() -> str:
    '''Compile the oracle library with its Makefile; returns its path.'''
    subprocess.run(['make', '-s', '-C', _here], check = True)
    return _libraryPath


_library = None


def library():
    '''The loaded oracle library, built first if absent.'''
    global _library
    if _library is None:
        if not os.path.exists(_libraryPath):
            build()
        _library = ctypes.CDLL(_libraryPath)
        _library.oracle_cfl_time_step.restype = ctypes.c_double
        _library.oracle_advance.restype = ctypes.c_int
        _library.oracle_advance_options.restype = ctypes.c_int
        _library.oracle_threads.restype = ctypes.c_int
    return _library


def _ptr(array: np.ndarray | None):
    if array is None:
        return None
    return array.ctypes.data_as(ctypes.c_void_p)


def _padded(values, count: int) -> np.ndarray:
    values = np.asarray(values, dtype = float).reshape(count, -1)
    result = np.zeros((count, 3))
    result[:, :values.shape[1]] = values
    return np.ascontiguousarray(result)


@dataclass
class Physics:

    '''The run's grid and physics, as the HIP library's run description carries them.'''

    dims: int
    h: float
    domainMin: np.ndarray
    domainMax: np.ndarray
    kernelType: str = 'wendlandC2'
    period: float = 0.0
    rho0: float = 1000.0
    c0: float = 20.0
    alpha: float = 0.1
    boundaryFactor: float = 0.0
    gravity: np.ndarray = field(default_factory = lambda: np.array([0.0, -9.81]))
    diffusion: str = 'none'
    delta: float = 0.1
    extrapolatedWalls: bool = False
    negativePressure: bool = False
    correction: bool = False
    limiter: bool = False
    densityFloor: float = 0.0
    cfl: float = 0.25
    maxTimeStep: float = 1.0
    hmin: float | None = None
    grid: tuple | None = None

    def params(self) -> OracleParams:
        from oracle.array_formulation import gridDimensions
        grid = self.grid or gridDimensions(self.domainMin, self.domainMax, 2.0 * self.h, self.dims)
        p = OracleParams()
        p.dims = self.dims
        p.kernel = 1 if self.kernelType == 'wendlandC2' else 0
        p.grid[:] = [int(c) for c in grid]
        p.cell = 2.0 * self.h
        p.hmax = self.h
        p.hmin = self.hmin if self.hmin is not None else self.h
        p.period = float(self.period or 0.0)
        p.rho0, p.c0, p.alpha, p.boundaryFactor = self.rho0, self.c0, self.alpha, self.boundaryFactor
        g = np.zeros(3)
        given = np.asarray(self.gravity, dtype = float)
        if len(given) == self.dims:
            g[:self.dims] = given
        else:
            g[self.dims - 1] = given[-1]
        p.gravity[:] = list(g)
        p.diffusion = {'none': 0, 'molteni': 1, 'hydrostatic': 2, 'antuono': 3}[self.diffusion]
        p.delta = self.delta
        p.extrapolatedWalls = int(self.extrapolatedWalls)
        p.negativePressure = int(self.negativePressure)
        p.correction = int(self.correction)
        p.limiter = int(self.limiter)
        p.densityFloor = self.densityFloor
        p.cfl, p.maxTimeStep = self.cfl, self.maxTimeStep
        return p


def _lengths(physics: Physics, count: int, lengths) -> np.ndarray:
    if lengths is None:
        return np.full(count, physics.h)
    return np.ascontiguousarray(np.asarray(lengths, dtype = float))


def cellKeys(physics: Physics, positions) -> np.ndarray:
    count = len(positions)
    keys = np.zeros(count, dtype = np.int64)
    p = physics.params()
    library().oracle_cell_keys(ctypes.byref(p), count, _ptr(_padded(positions, count)), _ptr(keys))
    return keys


def neighbors(physics: Physics, positions, lengths = None, precision: str = 'float64') -> tuple[np.ndarray, np.ndarray]:
    '''CSR rows (offsets, neighbors) of every particle's neighbor set, rows ascending.'''
    count = len(positions)
    p = physics.params()
    pos = _padded(positions, count)
    lens = _lengths(physics, count, lengths)
    function = library().oracle_neighbors if precision == 'float64' else library().oracle_neighbors32
    counts = np.zeros(count, dtype = np.int64)
    function(ctypes.byref(p), count, _ptr(pos), _ptr(lens), _ptr(counts), None, None)
    offsets = np.zeros(count + 1, dtype = np.int64)
    np.cumsum(counts, out = offsets[1:])
    rows = np.zeros(int(offsets[-1]), dtype = np.int32)
    function(ctypes.byref(p), count, _ptr(pos), _ptr(lens), None, _ptr(offsets), _ptr(rows))
    return offsets, rows


def _state(physics, positions, velocities, densities, masses, lengths, isFluid):
    count = len(positions)
    return (count, _padded(positions, count),
            _padded(velocities if velocities is not None else np.zeros((count, physics.dims)), count),
            np.ascontiguousarray(np.asarray(densities, dtype = float)),
            np.ascontiguousarray(np.asarray(masses, dtype = float)), _lengths(physics, count, lengths),
            np.ascontiguousarray(np.asarray(isFluid, dtype = bool).view(np.uint8)))


def wallPressure(physics: Physics, positions, densities, pressures, masses, isFluid, lengths = None):
    '''Boundary pressures from the fluid and the densities that carry them (sphGpu.py:804-854).'''
    count, pos, vel, dens, mass, lens, fluid = _state(physics, positions, None, densities, masses, lengths, isFluid)
    press = np.ascontiguousarray(np.asarray(pressures, dtype = float))
    outP, outRho = np.zeros(count), np.zeros(count)
    library().oracle_wall_pressure(ctypes.byref(physics.params()), count, _ptr(pos), _ptr(vel), _ptr(dens), _ptr(press),
                                   _ptr(mass), _ptr(lens), _ptr(fluid), _ptr(outP), _ptr(outRho))
    return outP, outRho


def velocityGradients(physics: Physics, positions, velocities, densities, masses, isFluid, lengths = None):
    count, pos, vel, dens, mass, lens, fluid = _state(physics, positions, velocities, densities, masses, lengths, isFluid)
    out = np.zeros((count, 9))
    library().oracle_velocity_gradients(ctypes.byref(physics.params()), count, _ptr(pos), _ptr(vel), _ptr(dens),
                                        _ptr(mass), _ptr(lens), _ptr(fluid), _ptr(out))
    dim = physics.dims
    return out.reshape(count, 3, 3)[:, :dim, :dim].copy()


def correctionMatrices(physics: Physics, positions, densities, masses, isFluid, lengths = None):
    count, pos, vel, dens, mass, lens, fluid = _state(physics, positions, None, densities, masses, lengths, isFluid)
    out = np.zeros((count, 9))
    library().oracle_correction(ctypes.byref(physics.params()), count, _ptr(pos), _ptr(dens), _ptr(mass), _ptr(lens),
                                _ptr(fluid), _ptr(out))
    dim = physics.dims
    return out.reshape(count, 3, 3)[:, :dim, :dim].copy()


def _padMatrices(matrices, count: int, fill: float) -> np.ndarray:
    out = np.zeros((count, 3, 3))
    dim = matrices.shape[1]
    out[:, :dim, :dim] = matrices
    if dim == 2:
        out[:, 2, 2] = fill
    return np.ascontiguousarray(out.reshape(count, 9))


def accelerations(physics: Physics, positions, velocities, densities, pressures, masses, isFluid, lengths = None,
                  gradients = None, closure = None):
    '''Accelerations of the fluid (sphGpu.py:303-414); `closure` adds the delta-LES-SPH viscosity.'''
    count, pos, vel, dens, mass, lens, fluid = _state(physics, positions, velocities, densities, masses, lengths, isFluid)
    press = np.ascontiguousarray(np.asarray(pressures, dtype = float))
    g = _padMatrices(gradients, count, 0.0) if gradients is not None else None
    c, keep = closureStruct(physics, closure)
    out = np.zeros((count, 3))
    library().oracle_accelerations(ctypes.byref(physics.params()), count, _ptr(pos), _ptr(vel), _ptr(dens), _ptr(press),
                                   _ptr(mass), _ptr(lens), _ptr(fluid), _ptr(g),
                                   ctypes.byref(c) if c is not None else None, _ptr(out))
    del keep
    return out[:, :physics.dims].copy()


def shiftingDisplacement(physics: Physics, positions, velocities, densities, masses, isFluid, size: float,
                         local: bool, spacing: float):
    '''Each particle's shift, density change and velocity change, on one h (sphGpu.py:856-1004).'''
    from oracle.array_formulation import kernelValues
    count, pos, vel, dens, mass, lens, fluid = _state(physics, positions, velocities, densities, masses, None, isFluid)
    spacingKernel = float(kernelValues(physics.kernelType, np.array([spacing]), physics.h, physics.dims)[0])
    shift, change, velocity = np.zeros((count, 3)), np.zeros(count), np.zeros((count, 3))
    library().oracle_shifting(ctypes.byref(physics.params()), count, _ptr(pos), _ptr(vel), _ptr(dens), _ptr(mass),
                              _ptr(fluid), ctypes.c_double(size), ctypes.c_int32(1 if local else 0),
                              ctypes.c_double(spacingKernel), _ptr(shift), _ptr(change), _ptr(velocity))
    dim = physics.dims
    return shift[:, :dim].copy(), change, velocity[:, :dim].copy()


def closureCoefficients(physics: Physics, gradients, isFluid, soundSpeed: float, lengths = None):
    '''Each particle's delta-LES-SPH alpha and delta from its G, (N, dim, dim) (sphGpu.py:699-723).'''
    count = len(gradients)
    g = _padMatrices(np.asarray(gradients, dtype = float), count, 0.0)
    lens = _lengths(physics, count, lengths)
    fluid = np.ascontiguousarray(np.asarray(isFluid, dtype = bool).view(np.uint8))
    alphas, deltas = np.zeros(count), np.zeros(count)
    library().oracle_closure_coefficients(ctypes.byref(physics.params()), count, _ptr(g), _ptr(lens), _ptr(fluid),
                                          ctypes.c_double(soundSpeed), _ptr(alphas), _ptr(deltas))
    return alphas, deltas


def continuityRate(physics: Physics, positions, velocities, densities, masses, isFluid, lengths = None,
                   corrections = None, densityGradients = None, closure = None):
    '''

    d rho / dt (sphGpu.py:416-546); 'antuono' diffusion takes the renormalized `densityGradients`,
    and `closure` each fluid pair's delta from the harmonic mean of the two particles'.

    '''
    count, pos, vel, dens, mass, lens, fluid = _state(physics, positions, velocities, densities, masses, lengths, isFluid)
    l = _padMatrices(corrections, count, 1.0) if corrections is not None else None
    g = _padded(densityGradients, count) if densityGradients is not None else None
    if physics.diffusion == 'antuono' and g is None:
        raise ValueError("The 'antuono' diffusion needs the density gradients.")
    c, keep = closureStruct(physics, closure)
    out = np.zeros(count)
    library().oracle_continuity(ctypes.byref(physics.params()), count, _ptr(pos), _ptr(vel), _ptr(dens), _ptr(mass),
                                _ptr(lens), _ptr(fluid), _ptr(l), _ptr(g), ctypes.byref(c) if c is not None else None,
                                _ptr(out))
    del keep
    return out


def densityGradients(physics: Physics, positions, densities, masses, isFluid, corrections, lengths = None):
    '''Each particle's renormalized density gradient with the given L, (N, dim) (sphGpu.py:548-591).'''
    count, pos, vel, dens, mass, lens, fluid = _state(physics, positions, None, densities, masses, lengths, isFluid)
    l = _padMatrices(corrections, count, 1.0)
    out = np.zeros((count, 3))
    library().oracle_density_gradients(ctypes.byref(physics.params()), count, _ptr(pos), _ptr(dens), _ptr(mass),
                                       _ptr(lens), _ptr(fluid), _ptr(l), _ptr(out))
    return out[:, :physics.dims].copy()


def shepardDensity(physics: Physics, positions, densities, masses, isFluid, lengths = None):
    '''The Shepard-filtered density of every particle (sphGpu.py:768-802).'''
    count, pos, vel, dens, mass, lens, fluid = _state(physics, positions, None, densities, masses, lengths, isFluid)
    out = np.zeros(count)
    library().oracle_shepard(ctypes.byref(physics.params()), count, _ptr(pos), _ptr(dens), _ptr(mass), _ptr(lens),
                             _ptr(out))
    return out


def xsphVelocity(physics: Physics, positions, velocities, densities, masses, isFluid, epsilon: float, lengths = None):
    '''The XSPH transport velocity of the fluid, the boundary's velocity unchanged (sphGpu.py:725-766).'''
    count, pos, vel, dens, mass, lens, fluid = _state(physics, positions, velocities, densities, masses, lengths, isFluid)
    out = np.zeros((count, 3))
    library().oracle_xsph(ctypes.byref(physics.params()), count, _ptr(pos), _ptr(vel), _ptr(dens), _ptr(mass),
                          _ptr(lens), _ptr(fluid), ctypes.c_double(epsilon), _ptr(out))
    return out[:, :physics.dims].copy()


def cflTimeStep(h, c0, vmax, amax, cfl, maxdt) -> float:
    return library().oracle_cfl_time_step(ctypes.c_double(h), ctypes.c_double(c0), ctypes.c_double(vmax),
                                          ctypes.c_double(amax), ctypes.c_double(cfl), ctypes.c_double(maxdt))


def boundaryStruct(description: dict | None) -> tuple[OracleBoundary, list]:
    '''An OracleBoundary from a boundary description dict; returns it with the arrays it points into.'''
    b = OracleBoundary()
    keep = []
    if not description:
        b.kind = 0
        return b, keep
    kind = description['kind']
    b.clearance = float(description.get('clearance', 0.0))
    if kind == 'container':
        b.kind = 1
        b.lower[:] = list(np.pad(np.asarray(description['lower'], dtype = float), (0, 3))[:3])
        b.upper[:] = list(np.pad(np.asarray(description['upper'], dtype = float), (0, 3))[:3])
        b.openTop = int(description.get('openTop', False))
    elif kind == 'bed':
        b.kind = 2
        bx = np.ascontiguousarray(np.asarray(description['bedX'], dtype = float))
        bz = np.ascontiguousarray(np.asarray(description['bedZ'], dtype = float))
        keep += [bx, bz]
        b.nBed = len(bx)
        b.bedX = bx.ctypes.data_as(_dp)
        b.bedZ = bz.ctypes.data_as(_dp)
        b.sideWalls = int(description.get('sideWalls', False))
        b.width = float(description.get('width', 0.0))
        b.offshoreWall = int(description.get('offshoreWall', False))
        b.wallX = float(description.get('wallX', 0.0))
        b.pistonSpeed = float(description.get('pistonSpeed', 0.0))
        b.endLimit = float(description['endLimit'])
    elif kind == 'heightField':
        b.kind = 3
        table = np.ascontiguousarray(np.asarray(description['table'], dtype = float))
        keep.append(table)
        b.fieldNy, b.fieldNx = table.shape
        b.table = table.ctypes.data_as(_dp)
        b.originX, b.originY = description['originX'], description['originY']
        b.stepX, b.stepY = description['stepX'], description['stepY']
        b.lowY, b.highY = description['lowY'], description['highY']
        b.offshoreWall = int(description.get('offshoreWall', False))
        b.lowX, b.highX = description['lowX'], description['highX']
    else:
        raise ValueError(f'Unknown boundary kind {kind!r}.')
    return b, keep


def enforce(physics: Physics, boundary: dict, positions, velocities, isFluid) -> tuple[np.ndarray, np.ndarray]:
    count = len(positions)
    pos, vel = _padded(positions, count), _padded(velocities, count)
    fluid = np.ascontiguousarray(np.asarray(isFluid, dtype = bool).view(np.uint8))
    b, keep = boundaryStruct(boundary)
    library().oracle_enforce(ctypes.byref(physics.params()), ctypes.byref(b), count, _ptr(pos), _ptr(vel), _ptr(fluid))
    return pos[:, :physics.dims].copy(), vel[:, :physics.dims].copy()


class OracleRun:

    """

    A run stepped by the float64 restatement, state held as padded (N, 3) arrays. `filterInterval`
    runs the Shepard filter at every step > 0 that it divides (counted from `steps`); `closure`
    ({'soundSpeed', 'referenceDensity', 'kinematicViscosity'}) runs the delta-LES-SPH closure;
    `ledger` keeps the energy budget's 14 device terms (`ledgerTerms`) over every advance.

    """

    def __init__(self, physics: Physics, boundary: dict | None, positions, velocities, densities, masses, isFluid,
                 lengths = None, time: float = 0.0, filterInterval: int = 0, closure: dict | None = None,
                 ledger: bool = False, steps: int = 0) -> None:
        self.physics = physics
        count, self.pos, self.vel, self.dens, self.mass, self.lens, self.fluid = _state(
            physics, positions, velocities, densities, masses, lengths, isFluid)
        self.dens = self.dens.copy()
        self.count = count
        self.press = np.zeros(count)
        self.acc = np.zeros((count, 3))
        self.time = ctypes.c_double(time)
        self.lastDt = ctypes.c_double(0.0)
        self.params = physics.params()
        self.boundary, self._keep = boundaryStruct(boundary)
        self.steps = steps
        self.options = None
        if filterInterval or closure is not None:
            self.options = OracleOptions(filterInterval = int(filterInterval), closure = int(closure is not None))
            if closure is not None:
                self.options.closureSoundSpeed = closure['soundSpeed']
                self.options.closureDensity = closure['referenceDensity']
                self.options.closureWater = 2.0 * (physics.dims + 2) * closure['kinematicViscosity'] / closure['soundSpeed']
        self.ledger = np.zeros(len(ledgerTerms)) if ledger else None

    def advance(self, untilTime: float, maxSteps: int, record: bool = False):
        if self.options is not None or self.ledger is not None or self.physics.diffusion == 'antuono':
            if record:
                raise ValueError('The step history is recorded without the options and the ledger only.')
            counter = ctypes.c_int64(self.steps)
            taken = library().oracle_advance_options(
                ctypes.byref(self.params), ctypes.byref(self.boundary),
                ctypes.byref(self.options) if self.options is not None else None, self.count, _ptr(self.pos),
                _ptr(self.vel), _ptr(self.dens), _ptr(self.press), _ptr(self.acc), _ptr(self.mass), _ptr(self.lens),
                _ptr(self.fluid), ctypes.byref(self.time), ctypes.c_double(untilTime), int(maxSteps),
                ctypes.byref(self.lastDt), ctypes.byref(counter), _ptr(self.ledger))
            self.steps += taken
            return taken
        history = np.zeros((maxSteps, 3)) if record else None
        taken = library().oracle_advance(ctypes.byref(self.params), ctypes.byref(self.boundary), self.count,
                                         _ptr(self.pos), _ptr(self.vel), _ptr(self.dens), _ptr(self.press),
                                         _ptr(self.acc), _ptr(self.mass), _ptr(self.lens), _ptr(self.fluid),
                                         ctypes.byref(self.time), ctypes.c_double(untilTime), int(maxSteps),
                                         ctypes.byref(self.lastDt), _ptr(history))
        self.steps += taken
        return (taken, history[:taken]) if record else taken

    def budget(self) -> dict:
        """The ledger's terms by name [J] (per unit depth in 2D)."""
        if self.ledger is None:
            raise ValueError('This run keeps no ledger.')
        return dict(zip(ledgerTerms, self.ledger.tolist()))

    def fluidEnergy(self) -> float:
        """K + P + E_C of the fluid [J]: sum m (v^2 / 2 - g . x + e(rho))."""
        fluid = self.fluid != 0
        g = np.zeros(3)
        g[:self.physics.dims] = np.asarray(self.physics.gravity, dtype = float)[:self.physics.dims]
        elastic = elasticEnergy(self.physics, self.dens)
        m = self.mass[fluid]
        return float(np.sum(m * (0.5 * np.sum(self.vel[fluid] ** 2, axis = 1) - self.pos[fluid] @ g + elastic[fluid])))

    @property
    def positions(self):
        return self.pos[:, :self.physics.dims]

    @property
    def velocities(self):
        return self.vel[:, :self.physics.dims]


def elasticEnergy(physics: Physics, densities) -> np.ndarray:
    """The restatement's elastic energy per unit mass at each density [J/kg]."""
    dens = np.ascontiguousarray(densities, dtype = float)
    out = np.zeros(len(dens))
    params = physics.params()
    library().oracle_elastic_energy(ctypes.byref(params), len(dens), _ptr(dens), _ptr(out))
    return out


def threads(requested: int = 0) -> int:
    '''Set (when positive) and return the oracle's OpenMP thread count.'''
    return library().oracle_threads(int(requested))
