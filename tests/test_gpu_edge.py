"""Edge cases of the single-world path on the GPU: a world without bodies, a world of planes
only, and bodies created between steps (World::add after World::run re-uploads the scene and
keeps the system IDs), checked against the oracle."""
import numpy as np
import pytest

import oracle as O
from pe_amd import World

pytestmark = pytest.mark.gpu


def test_world_without_bodies_steps():
    w = World()
    w.createMaterial("m", 1000.0, 0.0, 0.2, 0.2)
    w.run(3, 1e-3)
    assert w.getNumberOfContacts() == 0 and w.size == 0
    w.close()


def test_planes_only_world_steps_and_stays():
    # planes do not move; their normal is recomputed from the orientation every step
    # (HCSITS.h:4171-4176), which leaves pe's libm artefacts (e.g. 2.2e-16 in the x-plane's z):
    # the same bits as the oracle's
    w = World()
    w.createMaterial("m", 1000.0, 0.0, 0.2, 0.2)
    w.setGravity(0, 0, -9.81)
    ow = O.OracleWorld(gravity=(0, 0, -9.81))
    ow.create_material(1000.0, 0.0, 0.1, 0.1)
    for n, d in [((0, 0, 1), 0.0), ((1, 0, 0), 0.0), ((0.3, 0.4, 0.866), 0.5)]:
        w.createPlane(0, n, d, 0)
        ow.create_plane(n, d, 0)
    before = w.getState()
    w.run(5, 1e-3)
    ow.step(1e-3, 5)
    after, o = w.getState(), ow.get_state()
    assert w.getNumberOfContacts() == 0
    for k in ("pos", "q"):
        assert np.array_equal(before[k], after[k]), k
    for k in ("normal", "d"):
        assert np.array_equal(after[k], o[k]), k
    w.close()


def _separated_spheres(k0, k1):
    # spheres 1.5 m apart on a 4 x 4 lattice, each touching nothing but (later) the floor
    return [(1.0 + 1.5 * (k % 4), 1.0 + 1.5 * (k // 4), 0.4 + 0.05 * (k % 3)) for k in range(k0, k1)]


def test_bodies_added_between_steps_match_oracle():
    # Every sphere has at most one contact (with the floor) and starts at rest, so the contacts
    # are static and independent: the GPU's coloured sweeps and the oracle's pe-order sweeps
    # give the same bits.  Spheres created after 40 steps take the next system IDs.
    mu, dt = 0.4, 1e-3
    w = World()
    w.setGravity(0, 0, -9.81)
    w.setMaxIterations(10)
    w.setRelaxationModel(3)
    w.createMaterial("m", 2500.0, 0.0, mu / 2, mu / 2)
    ow = O.OracleWorld(gravity=(0, 0, -9.81), max_iterations=10, model=3)
    ow.create_material(2500.0, 0.0, mu / 2, mu / 2)
    w.createPlane(0, (0, 0, 1), 0.0, 0)
    ow.create_plane((0, 0, 1), 0.0, 0)
    for p in _separated_spheres(0, 8):
        w.createSphere(1, p, 0.4, 0)
        ow.create_sphere(p, 0.4, 0)
    w.run(40, dt)
    ow.step(dt, 40)
    for p in _separated_spheres(8, 16):
        w.createSphere(2, p, 0.4, 0)
        ow.create_sphere(p, 0.4, 0)
    w.run(60, dt)
    ow.step(dt, 60)
    g, o = w.getState(), ow.get_state()
    assert w.size == ow.n == 17
    assert w.getNumberOfContacts() == len(ow.contacts()["b1"]) > 0
    for k in ("pos", "v", "w", "q"):
        assert np.array_equal(g[k], o[k]), k
    assert np.array_equal(w.host_arrays()["sid"], ow.body_info()["sid"])
    w.close()
