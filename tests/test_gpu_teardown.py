"""Process exit with live library objects.

A program may end while a query server still has requests in flight, while
its Python wrappers sit in a reference cycle, or with a server handle it
never released.  The Python binding releases every handle it owns from an
interpreter-exit hook (reverse order of creation, before module teardown),
and the library stops the threads of servers still alive from a process-exit
handler that runs before the HIP runtime's own teardown.  The child process
below must exit with status 0 and print its last line.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import ctypes, sys
sys.path.insert(0, %r)
import openpgl_amd as pgl
from openpgl_amd.hip_runtime import DeviceArray
from openpgl_amd.synth import generate, queries
dev = pgl.Device()
f = pgl.Field(dev, max_samples_per_leaf=4000)
f.update(pgl.SampleStorage(generate(40000, seed=1)))
d = DeviceArray.from_numpy(generate(40000, seed=1, frame=1))
f.update_surface_device(d.value, 40000)      # queued, still running at exit
srv = pgl.Server(f)
pos, nrm, u2, vol, g = queries(60000, seed=2)
for i in range(6):
    srv.submit(pos, nrm, u2, g=g, is_volume=vol)   # in flight at exit
leaked = f.lib.pglNewServer(f.h, ctypes.byref(pgl.ServerConfig(0, 0, 0)))
assert leaked                                  # never released by the program
cycle = [srv, f, dev]
cycle.append(cycle)
print("last line", flush=True)
''' % ROOT


def test_exit_with_live_objects():
    print("child process", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD], capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.returncode
    assert "last line" in r.stdout
