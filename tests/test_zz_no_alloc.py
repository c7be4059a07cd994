"""pehip_step_noalloc, the time step that allocates nothing, against pehip_step.

The no-allocation variant keeps a step's per-contact arrays in the library's static device scratch
(at most 16,384 contacts); its results must equal the ordinary step's bit for bit: states, contact
sets and impulses, on two calls in a row (the second on a smaller scene, so that the scratch holds
the first call's larger contents), after switching back to the ordinary step, and a step that
detects more contacts than the scratch holds must be refused before it changes the state.
"""
import numpy as np
import pytest

from pe_amd import World, scenes, PeHipError, _abi

pytestmark = pytest.mark.gpu


def packed(nx, ny, nz, seed):
    # touching spheres on a cubic lattice resting on the floor, small random velocities:
    # about 3 contacts per sphere
    return scenes.drop_scene(nx, ny, nz, r=0.5, spacing=1.0 - 2e-6, jitter=2e-7,
                             box=(nx + 1.5, ny + 1.5, 4.0 * nz + 10.0), mu=0.4, seed=seed, v_amp=0.05,
                             max_iterations=25, z0=0.5 - 1e-6)


def make(sc, model=3):
    w = World()
    sc.to_world(w)
    w.setMaxIterations(sc.max_iterations)
    w.setRelaxationModel(model)
    w._ensure_device()
    return w


def step_noalloc(w, dt, n):
    _abi.check(_abi.load().pehip_step_noalloc(w._ctx, float(dt), int(n)), w._ctx)


def same(a, b):
    sa, sb = a.getState(), b.getState()
    for k in ("pos", "q", "v", "w"):
        assert np.array_equal(sa[k], sb[k]), k
    ca, cb = a.contacts(), b.contacts()
    for k in ca:
        assert np.array_equal(ca[k], cb[k]), k
    assert np.array_equal(a.impulses(), b.impulses())
    oa, ob = a.sweepOrder(), b.sweepOrder()
    assert np.array_equal(oa[0], ob[0]) and np.array_equal(oa[1], ob[1])
    return len(ca["b1"])


@pytest.mark.parametrize("model", [3, 1])
def test_noalloc_step_equals_step_on_two_calls(model):
    counts = []
    for sc in (packed(20, 20, 6, seed=3), packed(6, 6, 4, seed=4)):   # second call: smaller
        ref, na = make(sc, model), make(sc, model)
        ref.run(3, sc.dt)
        step_noalloc(na, sc.dt, 3)
        counts.append(same(ref, na))
        # back to the context's own buffers, then the static scratch again
        ref.run(1, sc.dt)
        na.run(1, sc.dt)
        same(ref, na)
        ref.run(1, sc.dt)
        step_noalloc(na, sc.dt, 1)
        same(ref, na)
        ref.close()
        na.close()
    assert counts[0] > 4 * counts[1] > 0 and counts[0] <= 16384, counts


def test_noalloc_step_refuses_more_contacts_than_the_scratch_holds():
    sc = packed(26, 26, 10, seed=5)                  # about 20,000 contacts
    w = make(sc)
    st0 = w.getState()
    with pytest.raises(PeHipError) as e:
        step_noalloc(w, sc.dt, 1)
    assert e.value.status == _abi.INVALID_ARG and "16384" in str(e.value)
    st1 = w.getState()
    for k in ("pos", "q", "v", "w"):
        assert np.array_equal(st0[k], st1[k]), k
    w.run(1, sc.dt)                                  # the ordinary step still takes it
    assert w.stats()["num_contacts"] > 16384
    w.close()
