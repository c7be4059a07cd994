"""Parity at full size: the 200,000-sphere polydisperse pile after 1,000 steps on the GPU, and
the 1,000,000-body mixed-shape hopper (950k spheres + 50k boxes) at the end of its 1,000-step job.

From the GPU's own state (identical inputs on both sides):
- the broadphase contact set equals the oracle's HashGrids detection bit for bit;
- a coloured step (50 sweeps) with the correctly rounded trig equals the oracle's sequential
  Gauss-Seidel in the GPU's colour order (oracle in correctly rounded trig mode) to 1e-12;
- the production step (50 sweeps, device libm, dataflow schedule) stays within 1e-4 (measured
  1.7e-5 on the angular velocities) of the oracle's Gauss-Seidel in the same order with the
  host libm: neither libm is correctly rounded, they differ in the last bit on some arguments,
  and such a difference can flip a golden-section comparison;
- the dataflow schedule and the colour launches give the same bits.
For the hopper the correctly rounded comparison runs three sweeps (the oracle's quad-precision
trig over ~10^5 contacts); the bar does not depend on the number of sweeps.
"""
import numpy as np
import pytest

import oracle as O
from pe_amd import World, scenes, SWEEP_COLOURED
from pe_amd._abi import SCHEDULE_COLOUR_LAUNCHES, SCHEDULE_DATAFLOW, SCHEDULE_ISLANDS

pytestmark = pytest.mark.gpu

STEPS = 1000


def rel_err(a, b, floor=1e-30):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), floor))


def contact_map(c):
    out = {}
    for i in range(len(c["b1"])):
        key = (int(c["b1"][i]), int(c["b2"][i]))
        out.setdefault(key, []).append((tuple(c["pos"][i]), tuple(c["normal"][i]), float(c["dist"][i]),
                                        float(c["mu"][i])))
    return out


def pair_ranks(c):
    seen = {}
    r = np.zeros(len(c["b1"]), np.uint32)
    for i, (a, b) in enumerate(zip(c["b1"], c["b2"])):
        k = (int(a), int(b))
        r[i] = seen.get(k, 0)
        seen[k] = int(r[i]) + 1
    return r


def gpu_world(sc, iters, exact=False, schedule=None):
    w = World()
    sc.to_world(w)
    w.setMaxIterations(iters)
    w.setRelaxationModel(sc.model)
    w.setSweepMode(SWEEP_COLOURED)
    w.setExactTrig(exact)
    if schedule is not None:
        w.setSweepSchedule(schedule)
    return w


def oracle_world(sc, iters, cr_trig):
    ow = O.OracleWorld(gravity=sc.gravity, max_iterations=iters, model=sc.model, cr_trig=cr_trig)
    sc.to_oracle(ow)
    return ow


@pytest.fixture(scope="module")
def c2_state():
    sc = scenes.config2()
    w = gpu_world(sc, sc.max_iterations)
    w.run(STEPS, sc.dt)
    st = w.getState()
    n = w.stats()["num_contacts"]
    w.close()
    print("C2 state after %d GPU steps: %d contacts" % (STEPS, n), flush=True)
    return sc, st


def one_gpu_step(sc, st, iters, exact, schedule=None):
    w = gpu_world(sc, iters, exact, schedule)
    w.setState(st)
    w.simulationStep(sc.dt)
    out = w.getState()
    idx, offs = w.sweepOrder()
    gc = w.contacts()
    stats = w.stats()
    w.close()
    return out, idx, offs, gc, stats


def oracle_step_in_order(sc, st, iters, cr_trig, gc, idx):
    ow = oracle_world(sc, iters, cr_trig)
    ow.set_state(st)
    ow.set_sweep_order_ranked(gc["b1"][idx], gc["b2"][idx], pair_ranks(gc)[idx])
    ow.step(sc.dt)
    return ow.get_state()


def test_c2_contact_set_bitexact(c2_state):
    sc, st = c2_state
    w = gpu_world(sc, sc.max_iterations)
    w.setState(st)
    g = contact_map(w.detect())
    w.close()
    ow = oracle_world(sc, sc.max_iterations, False)
    ow.set_state(st)
    hg = contact_map(ow.detect(exhaustive=False))
    assert sum(len(v) for v in hg.values()) > 1000
    assert g == hg


def test_c2_coloured_step_exact_trig_equals_gs_in_colour_order(c2_state):
    sc, st = c2_state
    g1, idx, offs, gc, _ = one_gpu_step(sc, st, sc.max_iterations, exact=True)
    assert len(gc["b1"]) > 1000 and offs[-1] == len(idx)
    o1 = oracle_step_in_order(sc, st, sc.max_iterations, True, gc, idx)
    for k in ("v", "w", "pos", "q"):
        assert rel_err(g1[k], o1[k]) <= 1e-12, k


def test_c2_production_step_close_to_gs_in_colour_order(c2_state):
    sc, st = c2_state
    g1, idx, _, gc, stats = one_gpu_step(sc, st, sc.max_iterations, exact=False)
    assert stats["schedule_used"] == SCHEDULE_DATAFLOW
    o1 = oracle_step_in_order(sc, st, sc.max_iterations, False, gc, idx)
    for k in ("v", "w"):
        assert rel_err(g1[k], o1[k]) <= 1e-4, k


@pytest.mark.parametrize("exact", [False, True])
def test_c2_dataflow_equals_colour_launches(c2_state, exact):
    sc, st = c2_state
    a = one_gpu_step(sc, st, sc.max_iterations, exact, SCHEDULE_DATAFLOW)
    b = one_gpu_step(sc, st, sc.max_iterations, exact, SCHEDULE_COLOUR_LAUNCHES)
    assert a[4]["schedule_used"] == SCHEDULE_DATAFLOW and b[4]["schedule_used"] == SCHEDULE_COLOUR_LAUNCHES
    for k in ("v", "w", "pos", "q"):
        assert np.array_equal(a[0][k], b[0][k]), k


@pytest.mark.parametrize("exact", [False, True])
def test_c2_islands_equal_colour_launches(c2_state, exact):
    # each island of the pile's contact graph swept on its own, colour by colour: the colour
    # launches' per-body update order, so the same bits
    sc, st = c2_state
    a = one_gpu_step(sc, st, sc.max_iterations, exact, SCHEDULE_ISLANDS)
    b = one_gpu_step(sc, st, sc.max_iterations, exact, SCHEDULE_COLOUR_LAUNCHES)
    assert a[4]["schedule_used"] == SCHEDULE_ISLANDS and b[4]["schedule_used"] == SCHEDULE_COLOUR_LAUNCHES
    for k in ("v", "w", "pos", "q"):
        assert np.array_equal(a[0][k], b[0][k]), k


def test_c2_islands_several_steps_equal_colour_launches(c2_state):
    # 20 steps in a row (every step re-finds and re-packs the islands)
    sc, st = c2_state
    out = []
    for schedule in (SCHEDULE_ISLANDS, SCHEDULE_COLOUR_LAUNCHES):
        w = gpu_world(sc, sc.max_iterations, False, schedule)
        w.setState(st)
        for _ in range(20):
            w.simulationStep(sc.dt)
            assert w.stats(kinetic=False)["schedule_used"] == schedule
        out.append((w.getState(), w.impulses()))
        w.close()
    for k in ("v", "w", "pos", "q"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert np.array_equal(out[0][1], out[1][1])


# --------------------------------------------------------------------------- 1M-body hopper
@pytest.fixture(scope="module")
def c3_state():
    sc = scenes.config3()
    w = gpu_world(sc, sc.max_iterations)
    w.run(sc.steps, sc.dt)
    st = w.getState()
    n = w.stats()["num_contacts"]
    w.close()
    print("C3 state after %d GPU steps: %d contacts" % (sc.steps, n), flush=True)
    return sc, st


def contact_kinds(sc, c):
    """Counts of contacts by shape pair (body order: planes, spheres, boxes)."""
    p, b = len(sc.planes), len(sc.planes) + sc.n_spheres
    b1, b2 = np.asarray(c["b1"]), np.asarray(c["b2"])
    kind = lambda i: np.where(i < p, 0, np.where(i < b, 1, 2))
    k1, k2 = kind(b1), kind(b2)
    names = {(1, 0): "sphere-plane", (2, 0): "box-plane", (1, 1): "sphere-sphere", (1, 2): "sphere-box",
             (2, 1): "sphere-box", (2, 2): "box-box"}
    out = {}
    for key, name in names.items():
        out[name] = out.get(name, 0) + int(((k1 == key[0]) & (k2 == key[1])).sum())
    return out


def same_contacts(g, o):
    """Equal contact sets: same records per body pair, in the same order within a pair."""
    if len(g["b1"]) != len(o["b1"]):
        return False
    def canon(c):
        key = (np.asarray(c["b1"], np.uint64) << np.uint64(32)) | np.asarray(c["b2"], np.uint64)
        i = np.argsort(key, kind="stable")
        return key[i], np.asarray(c["pos"])[i], np.asarray(c["normal"])[i], np.asarray(c["dist"])[i], \
            np.asarray(c["mu"])[i]
    return all(np.array_equal(a, b) for a, b in zip(canon(g), canon(o)))


def test_c3_contact_set_bitexact(c3_state):
    sc, st = c3_state
    w = gpu_world(sc, sc.max_iterations)
    w.setState(st)
    g = w.detect()
    w.close()
    ow = oracle_world(sc, sc.max_iterations, False)
    ow.set_state(st)
    o = ow.detect(exhaustive=False)
    print("C3 contacts by kind:", contact_kinds(sc, o), flush=True)
    assert len(o["b1"]) > 1000
    assert same_contacts(g, o)


def test_c3_compressed_contact_set_bitexact():
    # the hopper's initial lattice squeezed to 0.6 of its spacing about the channel axis and moved
    # 1 m into the left inclined plane: every shape pair in contact (2.35M contacts: box-box 73,
    # box-plane 2,307, sphere-box 186k, sphere-plane 27k), deep penetrations included (detection only)
    sc = scenes.config3()
    w = gpu_world(sc, sc.max_iterations)
    st = w.getState()
    p = len(sc.planes)
    n_left = np.array([0.5, 0.0, np.cos(np.radians(30.0))])
    st["pos"][p:] = (st["pos"][p:] - np.array([60.0, 60.0, 0.0])) * 0.6 + np.array([60.0, 60.0, 0.0]) - n_left
    w.setState(st)
    g = w.detect()
    w.close()
    ow = oracle_world(sc, sc.max_iterations, False)
    ow.set_state(st)
    o = ow.detect(exhaustive=False)
    kinds = contact_kinds(sc, o)
    print("compressed C3 contacts by kind:", kinds, flush=True)
    assert min(kinds.values()) > 50
    assert same_contacts(g, o)


def test_c3_coloured_step_exact_trig_equals_gs_in_colour_order(c3_state):
    sc, st = c3_state
    g1, idx, offs, gc, _ = one_gpu_step(sc, st, 3, exact=True)
    assert offs[-1] == len(idx) == len(gc["b1"])
    o1 = oracle_step_in_order(sc, st, 3, True, gc, idx)
    for k in ("v", "w", "pos", "q"):
        assert rel_err(g1[k], o1[k]) <= 1e-12, k


def test_c3_islands_equal_colour_launches(c3_state):
    sc, st = c3_state
    a = one_gpu_step(sc, st, sc.max_iterations, False, SCHEDULE_ISLANDS)
    b = one_gpu_step(sc, st, sc.max_iterations, False, SCHEDULE_COLOUR_LAUNCHES)
    assert a[4]["schedule_used"] == SCHEDULE_ISLANDS
    for k in ("v", "w", "pos", "q"):
        assert np.array_equal(a[0][k], b[0][k]), k


def test_c3_production_step_close_to_gs_in_colour_order(c3_state):
    sc, st = c3_state
    g1, idx, _, gc, _ = one_gpu_step(sc, st, sc.max_iterations, exact=False)
    o1 = oracle_step_in_order(sc, st, sc.max_iterations, False, gc, idx)
    for k in ("v", "w"):
        assert rel_err(g1[k], o1[k]) <= 1e-4, k
