"""Edge cases of the single-world path on the GPU: a world without bodies, a world of planes
only, and bodies created between steps (World::add after World::run re-uploads the scene and
keeps the system IDs), checked against the oracle; material tables that grow between steps and
refuse to shrink below the uploaded bodies' material IDs."""
import numpy as np
import pytest

import oracle as O
from pe_amd import World, scenes, _abi
from pe_amd.batch import BatchWorld

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


def _resting_lattice(seed):
    # touching spheres on a cubic lattice resting on the floor, small random velocities
    return scenes.drop_scene(4, 4, 2, r=0.5, spacing=1.0 - 2e-6, jitter=2e-7, box=(5.5, 5.5, 18.0), mu=0.4,
                             seed=seed, v_amp=0.05, max_iterations=25, z0=0.5 - 1e-6)


def _grown(t, nm):
    g = np.zeros((nm + 1, nm + 1))
    g[:nm, :nm] = np.asarray(t, float).reshape(nm, nm)
    return g


def test_material_tables_grow_and_never_shrink_below_the_bodies():
    # The bodies index the friction tables by material ID (row stride = the table's size), so a
    # table smaller than the one the bodies were uploaded with is refused: it would be read past
    # its end.  A larger table whose leading block is the old one is taken, and the world then
    # steps bit for bit as one that kept the old table.
    L = _abi.load()
    sc = _resting_lattice(seed=5)
    ref, w = World(), World()
    for x in (ref, w):
        sc.to_world(x)
        x.run(3, sc.dt)
    nm = len(w.materials)
    csf, cdf = _grown(w._csf, nm), _grown(w._cdf, nm)
    _abi.check(L.pehip_set_materials(w._ctx, nm + 1, _abi.ptr(csf), _abi.ptr(cdf)), w._ctx)
    small = np.ascontiguousarray(np.asarray(w._csf, float).reshape(nm, nm))
    st = L.pehip_set_materials(w._ctx, nm, _abi.ptr(small), _abi.ptr(small))
    assert st == _abi.INVALID_ARG and "shrink" in L.pehip_last_error(w._ctx).decode()
    ref.run(5, sc.dt)
    w.run(5, sc.dt)
    assert w.getNumberOfContacts() > 20
    a, b = ref.getState(), w.getState()
    for k in ("pos", "q", "v", "w"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(ref.impulses(), w.impulses())
    ref.close()
    w.close()

    # the batched scenes follow the same rule
    scs = [_resting_lattice(seed=6), _resting_lattice(seed=7)]
    bref, bw = BatchWorld(scs), BatchWorld(scs)
    nm = len(scs[0].materials)
    csf = np.zeros((nm, nm))
    for i, mi in enumerate(scs[0].materials):
        for j, mj in enumerate(scs[0].materials):
            csf[i, j] = mi["csf"] + mj["csf"]
    cdf = np.zeros((nm, nm))
    for i, mi in enumerate(scs[0].materials):
        for j, mj in enumerate(scs[0].materials):
            cdf[i, j] = mi["cdf"] + mj["cdf"]
    g1, g2 = _grown(csf, nm), _grown(cdf, nm)
    _abi.check_batch(L.pehip_batch_set_materials(bw._b, nm + 1, _abi.ptr(g1), _abi.ptr(g2)), bw._b)
    st = L.pehip_batch_set_materials(bw._b, nm, _abi.ptr(csf), _abi.ptr(cdf))
    assert st == _abi.INVALID_ARG and "shrink" in L.pehip_batch_last_error(bw._b).decode()
    bref.run(5)
    bw.run(5)
    a, b = bref.getState(), bw.getState()
    for k in ("pos", "q", "v", "w"):
        assert np.array_equal(a[k], b[k]), k
    assert len(bw.contacts(0)["b1"]) > 20
    bref.close()
    bw.close()
