"""Forced failures of the legacy (value-returning) entry points, on the CPU: like the
reference, whose C++ exceptions terminate the process, they report the error on stderr and
abort. Each case runs in a child process, before any device work:

* a radial transform of fewer than two points (src/horton/rtransform.cpp:38);
* two centres at the same position in a Becke grid (the cell functions divide by every
  centre distance, src/grid/becke.cpp:150), checked before the grid touches the device."""
import os
import signal
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _child(code):
    env = dict(os.environ, PYTHONPATH=ROOT)
    return subprocess.run([sys.executable, '-c', code], cwd=ROOT, env=env, capture_output=True,
                          text=True, timeout=120)


def _library_built():
    return os.path.exists(os.path.join(ROOT, 'napmo_amd', 'libnapmo_gpu.so'))


@pytest.mark.skipif(not _library_built(), reason='libnapmo_gpu.so not built')
def test_radial_transform_of_one_point_aborts_with_message():
    r = _child('from napmo_amd.system import cext\n'
               'cext.ChebyshevRTransform_new(1.0, 1)\n'
               'print("returned")')
    assert r.returncode == -signal.SIGABRT, (r.returncode, r.stdout, r.stderr)
    assert 'ChebyshevRTransform_new' in r.stderr
    assert 'at least two points' in r.stderr
    assert 'returned' not in r.stdout


@pytest.mark.skipif(not _library_built(), reason='libnapmo_gpu.so not built')
def test_coincident_centres_abort_before_device_work():
    r = _child('import napmo_amd\n'
               'atoms = [("H", (0.0, 0.0, 0.7)), ("H", (0.0, 0.0, 0.7))]\n'
               'napmo_amd.BeckeGrid(napmo_amd.make_species(atoms), 20, 110)\n'
               'print("returned")')
    assert r.returncode == -signal.SIGABRT, (r.returncode, r.stdout, r.stderr)
    assert 'coincident centres in BeckeGrid (0 and 1)' in r.stderr
    assert 'no HIP device' not in r.stderr
    assert 'returned' not in r.stdout
