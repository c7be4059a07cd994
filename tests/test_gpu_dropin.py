"""The C++ drop-in API (include/pe_amd/pe.hpp) driven like a pe program: its result equals
the Python World's on the same scene bit for bit (same kernels, same inputs)."""
import os
import subprocess

import numpy as np
import pytest

from pe_amd import World
from pe_amd._abi import SCHEDULE_COLOUR_LAUNCHES, SCHEDULE_DATAFLOW

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pe_amd", "_lib", "granular_dropin")


def python_run(nx, ny, nz, steps, ncaps=0, nunions=0):
    w = World()
    w.setGravity(0, 0, -9.81)
    w.setMaxIterations(50)
    w.setRelaxationModel(3)
    m = w.createMaterial("granular", 2500, 0, 0.2, 0.2)
    lx, ly, lz = 2.0 + nx * 1.2, 2.0 + ny * 1.2, 40.0
    for n, d in [((0, 0, 1), 0.0), ((0, 0, -1), -lz), ((0, 1, 0), 0.0), ((0, -1, 0), -ly), ((1, 0, 0), 0.0),
                 ((-1, 0, 0), -lx)]:
        w.createPlane(0, n, d, m)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                jit = 0.001 * float((i * 7 + j * 13 + k * 17) % 11 - 5)
                w.createSphere(6, (1.5 + 1.2 * i + jit, 1.5 + 1.2 * j - jit, 0.6 + 1.2 * k + jit), 0.5, m)
    for c in range(ncaps):
        w.createCapsule(7, (1.6 + 1.1 * c, 0.5 * ly + 0.05 * c, 0.6 + 1.2 * (nz - 1) + 0.77), 0.25, 0.8, m)
    for u in range(nunions):
        w.createUnion(8, (1.7 + 1.0 * u, 0.5 * ly - 0.6 + 0.07 * u, 0.6 + 1.2 * (nz - 1) + 1.9),
                      [((-0.25, 0.0, 0.02), 0.22, m), ((0.25, 0.03, 0.0), 0.2, m), ((0.0, 0.2, -0.1 * (u % 2)), 0.18, m)])
    w.run(steps, 1e-3)
    return w


@pytest.mark.parametrize("ncaps,nunions", [(0, 0), (3, 0), (2, 3)])
def test_cpp_dropin_matches_python_world(ncaps, nunions):
    # with capsules and unions: pe::createCapsule's and pe::createUnion's properties equal the
    # Python World's bit for bit
    if not os.path.exists(EXE):
        pytest.fail("drop-in example not built")
    nx, ny, nz, steps = 5, 4, 3, 220
    r = subprocess.run([EXE, str(nx), str(ny), str(nz), str(steps), str(ncaps), str(nunions)], capture_output=True, text=True,
                       timeout=120)
    print(r.stderr)
    assert r.returncode == 0, r.stderr
    out = r.stdout.split("\n")
    head = out[0].split()
    rows = [l.split() for l in out[1:] if l.strip()]
    pos = np.array([[float.fromhex(x) for x in r[1:4]] for r in rows])
    vel = np.array([[float.fromhex(x) for x in r[4:7]] for r in rows])
    w = python_run(nx, ny, nz, steps, ncaps, nunions)
    st = w.getState()
    assert len(rows) == w.size
    assert int(head[1]) == w.getNumberOfContacts() and int(head[1]) > 0
    assert np.array_equal(pos, st["pos"]) and np.array_equal(vel, st["v"])
    # The automatic schedule: the dataflow sweeps, unless the device refuses the child process's
    # cooperative launch (observed while this test process holds cooperative launches of its own),
    # in which case the library says so once on stderr and sweeps with colour launches -- the same
    # bits, as the comparison above shows for whichever ran.
    refused = "dataflow sweep launch refused" in r.stderr
    assert int(head[5]) == (SCHEDULE_COLOUR_LAUNCHES if refused else SCHEDULE_DATAFLOW), (head, r.stderr)
    print("child schedule:", "colour launches (cooperative launch refused)" if refused else "dataflow")
