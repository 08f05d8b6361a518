'''

The float64 restatement under AddressSanitizer and UndefinedBehaviorSanitizer: `make -C oracle
asan` builds oracle/asan_check.cpp with the restatement and runs every entry point (2D and periodic
3D clouds, mixed lengths, a lone particle, ten steps in a box). A sanitizer finding aborts it.

'''

import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'oracle')


@pytest.mark.timeout(300)
def testTheRestatementRunsCleanUnderTheSanitizers():
    if shutil.which('g++') is None:
        pytest.skip('no g++ to build the sanitized restatement')
    result = subprocess.run(['make', '-s', '-C', ORACLE, 'asan'], capture_output = True, text = True)
    assert result.returncode == 0, result.stdout + result.stderr
    assert 'asan check passed: 10 steps' in result.stdout
    assert 'ERROR: AddressSanitizer' not in result.stderr and 'runtime error' not in result.stderr
