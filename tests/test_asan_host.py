"""Address and undefined-behaviour sanitizer runs of the host code, on the CPU: the library's
host routines (napmo_amd/csrc/host.cpp: transforms, Lebedev rules, radial and atomic grids,
splines and extrapolations, the Gaussian basis, harmonics) and the oracle's C restatement
(oracle/napmo_oracle.c), each built with -fsanitize=address,undefined into a small driver
under tests/asan/ and run; any sanitizer report fails the test. No device code is built."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ['-O1', '-g', '-fno-omit-frame-pointer', '-fsanitize=address,undefined',
       '-fno-sanitize-recover=undefined']


def _run(cmd, exe, tmp_path):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS='abort_on_error=0:halt_on_error=1',
               UBSAN_OPTIONS='print_stacktrace=1:halt_on_error=1')
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env, cwd=tmp_path)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-3000:]
    assert 'ERROR: AddressSanitizer' not in r.stderr and 'runtime error' not in r.stderr
    return r.stdout


@pytest.mark.skipif(shutil.which('g++') is None or not os.path.isdir('/opt/rocm/include'),
                    reason='g++ or the HIP headers are absent')
def test_host_code_under_address_sanitizer(tmp_path):
    exe = str(tmp_path / 'host_asan')
    gen = os.path.join(ROOT, 'napmo_amd', 'csrc', 'lebedev_table.inc')
    if not os.path.exists(gen):
        pytest.skip('lebedev_table.inc not generated (run napmo_amd/build.py)')
    cmd = (['g++', '-std=c++17'] + SAN + ['-D__HIP_PLATFORM_AMD__', '-I/opt/rocm/include',
           '-I' + os.path.join(ROOT, 'napmo_amd', 'csrc'), '-I' + os.path.join(ROOT, 'include'),
           os.path.join(ROOT, 'napmo_amd', 'csrc', 'host.cpp'), os.path.join(ROOT, 'tests', 'asan', 'host_asan.cpp'),
           '-o', exe])
    assert 'host asan driver: ok' in _run(cmd, exe, tmp_path)


@pytest.mark.skipif(shutil.which('gcc') is None, reason='gcc is absent')
def test_oracle_c_under_address_sanitizer(tmp_path):
    exe = str(tmp_path / 'oracle_asan')
    cmd = (['gcc', '-std=gnu11', '-fopenmp'] + SAN + [os.path.join(ROOT, 'oracle', 'napmo_oracle.c'),
           os.path.join(ROOT, 'tests', 'asan', 'oracle_asan.c'), '-o', exe, '-lm'])
    assert 'oracle asan driver: ok' in _run(cmd, exe, tmp_path)
