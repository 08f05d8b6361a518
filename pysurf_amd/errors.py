'''The engine's exceptions, after the reference's errors module (src/pysurf/errors.py).'''


class EngineError(RuntimeError):
    '''The HIP library is missing, no HIP device is present, or a HIP call failed.'''


class NumericalInstabilityError(RuntimeError):
    '''A run's time step collapsed or its state stopped being finite (BreakingWave.py:894-903, 916-919).'''
