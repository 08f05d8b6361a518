'''

Many runs in one engine: every run's particles packed into consecutive slots of one context, each
with its own cell grid, physics, boundary and clock, all stepped by the same kernel launches and,
with `graph`, replayed from one captured HIP graph. The reference runs one `PairSums` per solver
and one run per process (sphSolver.py:1499-1502); packing is what fills an MI355X with 2D runs of
a few thousand to a hundred thousand particles.

A run's results do not depend on its neighbors in the batch: every sum is a gather over the run's
own cells, in the same order as when the run is alone, and each run's time step is its own.

'''

from __future__ import annotations

import numpy as np

from pysurf_amd import _native
from pysurf_amd.pair_sums import Engine, boundaryOf, gridDimensions
from pysurf_amd.particles import ParticleSystem

diffusionModes = {'none': 0, 'molteni': 1, 'hydrostatic': 2, 'antuono': 3}


def runOf(case, first: int, boundary: _native.Boundary) -> _native.Run:
    '''The sph_run of a scenarios.Case whose particles start at slot `first`.'''
    run = _native.Run()
    run.first, run.count = int(first), int(case.count)
    run.kernel = 1 if case.kernelType == 'wendlandC2' else 0
    grid = gridDimensions(case.domainMin, case.domainMax, 2.0 * case.smoothingLength, case.dimensions)
    run.grid[:] = list(grid)
    run.cell = 2.0 * case.smoothingLength
    lengths = case.particles.lengths
    run.hmax = case.smoothingLength
    run.hmin = float(lengths.min()) if lengths is not None else case.smoothingLength
    run.period = 0.0
    run.rho0, run.c0, run.alpha = case.referenceDensity, case.soundSpeed, case.alpha
    run.boundaryFactor = case.boundaryViscosityFactor
    g = np.zeros(3)
    g[:case.dimensions] = case.gravity
    run.gravity[:] = list(g)
    run.diffusion = diffusionModes[case.diffusion]
    run.delta = case.delta
    run.extrapolatedWalls = int(case.extrapolatedWalls)
    run.negativePressure = int(case.negativePressure)
    run.correction = int(case.correction)
    run.limiter = int(case.limiter)
    run.densityFloor = case.densityFloor
    run.cfl, run.maxTimeStep = case.cfl, case.maxTimeStep
    run.boundary = boundary
    run.filterInterval = int(getattr(case, 'filterInterval', 0))
    return run


# The state fields that make a run's state between steps, the ones BatchEngine's carry takes
stateFields = ('position', 'velocity', 'density', 'mass', 'length', 'fluid')


class BatchEngine:

    '''

    Runs (scenarios.Case) packed into one context of `precision` pair arithmetic.

    Parameters:
    -----------
    cases : list
        The runs, all of one dimension
    precision : str
        'float32' pairs over float64 state, or 'float64'
    device : str | None
        'cuda:N'; None for cuda:0
    blockSize : int
        Threads per workgroup of the sweeps
    lanes : int
        Lanes per particle of the sweeps; 0 picks by dimension
    listCapacity : int
        Neighbor-list entries per particle; 0 walks the cells in every sweep, -1 picks by dimension
    tune : bool
        Time the candidate launch shapes on these runs and keep the fastest (autotune.tune); the
        three arguments above are then ignored
    closure : dict | None
        The delta-LES-SPH closure of every run: {'soundSpeed', 'referenceDensity',
        'kinematicViscosity'} (PairSums.setClosure); None for none
    carry : list | None
        Per case, None to load its particles from the host, or (fields, start) to take its state from
        device tensors: `fields` maps each state field ('position', 'velocity', 'density', 'mass',
        'length', 'fluid') to a tensor laid out as the engine's (rows by slots) and the run's slots
        start at `start` there (the request loop's membership changes: the runs that stay never
        cross PCIe)

    '''

    def __init__(self, cases: list, precision: str = 'float32', device = None, blockSize: int = 256,
                 lanes: int = 0, listCapacity: int = -1, tune: bool = False, closure: dict | None = None,
                 carry: list | None = None) -> None:
        if not cases:
            raise ValueError('A batch needs at least one run.')
        dimensions = {case.dimensions for case in cases}
        if len(dimensions) != 1:
            raise ValueError('Every run of a batch has the same dimensions.')
        self.dimensions = dimensions.pop()
        self.cases = list(cases)
        self.engine = Engine(self.dimensions, precision, device)
        self.device = self.engine.device
        self.precision = precision
        self._keep: list = []
        runs, first = [], 0
        self.offsets = []
        for case in self.cases:
            boundary, keep = boundaryOf(case.boundary, self.dimensions, self.device)
            self._keep += keep
            runs.append(runOf(case, first, boundary))
            self.offsets.append(first)
            first += case.count
        self.offsets.append(first)
        self.fluidCounts = [case.fluidCount for case in self.cases]
        self.engine.configure(runs)
        if closure is not None:
            self.engine.setClosure(True, closure['soundSpeed'], closure['referenceDensity'],
                                   closure['kinematicViscosity'])
        self.engine.setLaunch(blockSize, lanes, listCapacity)
        self.identity = None
        self.load(carry)
        self.launch = {'lanes': lanes, 'block': blockSize, 'listCapacity': listCapacity, 'nearCells': -1,
                       'nearCellsBuild': -1}
        if tune:
            from pysurf_amd import autotune
            self.launch = autotune.tune(self)

    @property
    def count(self) -> int:
        return self.offsets[-1]

    def setEnergyBudget(self, enabled: bool) -> None:
        '''Every run's per-particle energy ledger on, from zero, or off (PairSums.setEnergyBudget).'''
        self.engine.setEnergyBudget(enabled)

    def energySums(self) -> list:
        '''Each run's energy budget terms since the ledger went on [J], a dict per run (PairSums.energySums).'''
        sums = self.engine.energySums(len(self.cases))
        return [dict(zip(_native.energyTermNames, map(float, row))) for row in sums]

    @property
    def fluidCount(self) -> int:
        return sum(self.fluidCounts)

    def load(self, carry: list | None = None) -> None:
        '''Copy every run's particle state into the engine, in the cases' own order: from the host, or
        for a run that `carry` names, from another engine's slots on the device.'''
        engine = self.engine
        self.identity = None
        carry = carry or [None] * len(self.cases)
        if len(carry) != len(self.cases):
            raise ValueError('The carried runs need one entry per case.')
        fields = stateFields
        hosts = [index for index, source in enumerate(carry) if source is None]
        if len(hosts) == len(self.cases):
            parts = [case.particles for case in self.cases]
            engine.upload('position', np.vstack([p.positions for p in parts]))
            engine.upload('velocity', np.vstack([p.velocities for p in parts]))
            engine.upload('density', np.concatenate([p.densities for p in parts]))
            engine.upload('mass', np.concatenate([p.masses for p in parts]))
            lengths = [p.lengths if p.lengths is not None else np.full(p.nParticles, case.smoothingLength)
                       for p, case in zip(parts, self.cases)]
            engine.upload('length', np.concatenate(lengths))
            engine.upload('fluid', np.concatenate([p.isFluid for p in parts]).astype(np.uint8))
        else:
            for index, source in enumerate(carry):
                first = self.offsets[index]
                if source is None:
                    p, case = self.cases[index].particles, self.cases[index]
                    lengths = p.lengths if p.lengths is not None else np.full(p.nParticles, case.smoothingLength)
                    for name, values in (('position', p.positions), ('velocity', p.velocities),
                                         ('density', p.densities), ('mass', p.masses), ('length', lengths),
                                         ('fluid', p.isFluid.astype(np.uint8))):
                        engine.uploadSlice(name, values, first)
                    continue
                tensors, start = source
                count = self.cases[index].count
                for name in fields:
                    if tensors[name].shape[-1] < start + count:
                        raise ValueError(f'The carried {name} holds fewer than the run\'s {count} slots.')
                    engine.copySlice(name, tensors[name], start, first, count)
        engine.synchronize()
        # runs whose masses, lengths and kinds the device changed (a split or a merge)
        self.refined = [False] * len(self.cases)
        self.until = [0.0] * len(self.cases)
        self.maxSteps = [0] * len(self.cases)
        for r in range(len(self.cases)):
            self.setClock(r, 0.0, 0.0, 0, 0)

    def setClock(self, run: int, time: float, untilTime: float, steps: int, maxSteps: int) -> None:
        '''Run `run` is at `time` after `steps` steps, and steps while time < `untilTime` and steps < `maxSteps`.'''
        self.engine.setClock(run, time, untilTime, steps, maxSteps)
        self.until[run], self.maxSteps[run] = float(untilTime), int(maxSteps)

    def setClocks(self, untilTime: float = np.inf, maxSteps: int = 2**62, times = None, steps = None) -> None:
        '''Every run steps until its time reaches `untilTime` or it has taken `maxSteps` steps.'''
        for r in range(len(self.cases)):
            self.setClock(r, 0.0 if times is None else times[r], untilTime, 0 if steps is None else steps[r], maxSteps)

    def advance(self, steps: int, graph: bool = True) -> None:
        '''Enqueue `steps` steps of every active run (asynchronous).'''
        self.engine.advance(steps, graph)

    def synchronize(self) -> None:
        self.engine.synchronize()

    def clocks(self) -> dict:
        return self.engine.clock(len(self.cases))

    def sortParticles(self) -> None:

        '''

        Permute every run's particles into the cell order of their current positions, on the device
        (WcsphSolver.sortParticles, sphSolver.py:2093-2126): the gathers of the next steps then
        read neighboring slots. The engine remembers each slot's particle, so `state` still returns
        particles in the cases' order. Asynchronous.

        '''

        import torch
        if self.identity is None:
            with torch.cuda.device(self.device):
                self.identity = torch.arange(max(self.count, 1), dtype = torch.int32, device = self.device)
                torch.cuda.current_stream(self.device).synchronize()
        self.engine.sortState(self.identity)

    def _original(self, values: np.ndarray, a: int, b: int) -> np.ndarray:
        '''Slots a..b of a downloaded field, back in the case's particle order.'''
        if self.identity is None:
            return values[a:b]
        identity = self.identity[a:b].cpu().numpy().astype(np.int64) - a
        ordered = np.empty_like(values[a:b])
        ordered[identity] = values[a:b]
        return ordered

    def split(self, zones: list, lengthRatio: float = 1.0, minMass = 0.0, mask = None) -> np.ndarray:

        '''

        Split, on the device, each run's fluid particles with x in its zone (`zones`: one
        (lowX, highX) or None per run) and mass at least `minMass` (one value, or one per run; the
        cases' unsplit mass times 0.75 takes only particles that have not split), into 2^dims
        daughters whose smoothing length is `lengthRatio` their mother's (splitParticles,
        sphRefinement.py:135-186). Each run grows in place: its own particles keep their slots
        relative to its start, the other daughters follow them, mother by mother. The clocks carry
        over. Returns how many particles of each run split. The batch must be in slot order (not
        sorted). `load` afterwards restores the cases' own particles, and a checkpoint of the grown
        runs restores only into a batch of the same grown runs.

        '''

        if self.identity is not None:
            raise ValueError('Split a batch before sorting it: its slots must hold the cases\' own order.')
        runs = len(self.cases)
        if len(zones) != runs:
            raise ValueError(f'The split needs one zone (or None) per run, {runs} of them.')
        masses = np.broadcast_to(np.asarray(minMass, dtype = float), (runs,))
        specs = []
        for zone, least in zip(zones, masses):
            spec = _native.SplitSpec()
            if zone is not None:
                spec.active, spec.lowX, spec.highX, spec.minMass = 1, float(zone[0]), float(zone[1]), float(least)
            specs.append(spec)
        clock = self.clocks()
        chosen = self.engine.split(specs, self.cases[0].referenceDensity, lengthRatio, mask)
        self.refined = [refined or bool(count) for refined, count in zip(self.refined, chosen)]
        if chosen.any():
            grown = chosen * (2**self.dimensions - 1)
            counts = np.diff(self.offsets) + grown
            self.offsets = [0] + np.cumsum(counts).astype(int).tolist()
            self.fluidCounts = [count + int(extra) for count, extra in zip(self.fluidCounts, grown)]
            for r in range(runs):
                self.setClock(r, float(clock['time'][r]), self.until[r], int(clock['steps'][r]), self.maxSteps[r])
        return chosen

    def merge(self, zones: list, referenceMass, lengthRatio: float = 1.0, mask = None):

        '''

        Merge, on the device, each run's split fluid particles with x in its zone (`zones`: one
        (lowX, highX) or None per run), 2^dims of one level at a time into one particle of the level
        above (mergeParticles, sphRefinement.py:212-311), `referenceMass` (one value, or one per run)
        being an unsplit fluid particle's mass and `lengthRatio` the ratio the daughters' lengths
        took. Each run closes up in place; the clocks carry over. Returns a MergeResult per run
        (its `source` left empty). The batch must be in slot order (not sorted).

        '''

        from pysurf_amd.refinement import MergeResult
        if self.identity is not None:
            raise ValueError('Merge a batch before sorting it: its slots must hold the cases\' own order.')
        runs = len(self.cases)
        if len(zones) != runs:
            raise ValueError(f'The merge needs one zone (or None) per run, {runs} of them.')
        masses = np.broadcast_to(np.asarray(referenceMass, dtype = float), (runs,))
        specs = []
        for zone, mass, case in zip(zones, masses, self.cases):
            spec = _native.MergeSpec()
            if zone is not None:
                spec.active, spec.lowX, spec.highX = 1, float(zone[0]), float(zone[1])
                spec.referenceMass, spec.referenceDensity = float(mass), float(case.referenceDensity)
                spec.lengthRatio = float(lengthRatio)
            specs.append(spec)
        clock = self.clocks()
        groups, energy = self.engine.merge(specs, mask)
        self.refined = [refined or bool(count) for refined, count in zip(self.refined, groups)]
        if groups.any():
            shrunk = groups * (2**self.dimensions - 1)
            counts = np.diff(self.offsets) - shrunk
            self.offsets = [0] + np.cumsum(counts).astype(int).tolist()
            self.fluidCounts = [count - int(less) for count, less in zip(self.fluidCounts, shrunk)]
            for r in range(runs):
                self.setClock(r, float(clock['time'][r]), self.until[r], int(clock['steps'][r]), self.maxSteps[r])
        return [MergeResult(groups = int(g), removedEnergy = float(e)) for g, e in zip(groups, energy)]

    def globalSums(self) -> np.ndarray:
        '''Every run's float64 sums over its fluid, shape (runs, 7), columns as _native.globalSumNames.'''
        return self.engine.globalSums(len(self.cases))

    def state(self, index: int) -> ParticleSystem:
        '''Run `index`'s state as a ParticleSystem, copied from the device.'''
        a, b = self.offsets[index], self.offsets[index + 1]
        engine, d = self.engine, self.dimensions
        particles = self.cases[index].particles.copy()
        particles.positions = self._original(engine.download('position', d, b), a, b)
        particles.velocities = self._original(engine.download('velocity', d, b), a, b)
        particles.densities = self._original(engine.download('density', None, b), a, b)
        particles.pressures = self._original(engine.download('pressure', None, b), a, b)
        particles.accelerations = self._original(engine.download('acceleration', d, b), a, b)
        if self.refined[index] or b - a != self.cases[index].count:
            # The run has split or merged on the device (a split then a merge may restore its count):
            # its masses, lengths and kinds come from there too
            particles.masses = self._original(engine.download('mass', None, b), a, b)
            particles.lengths = self._original(engine.download('length', None, b), a, b)
            particles.isFluid = self._original(engine.download('fluid', None, b), a, b).astype(bool)
        return particles
