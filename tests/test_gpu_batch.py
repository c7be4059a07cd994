"""Batched scenes on the GPU: every scene of a batch obeys the bars of the single-world path.

Each scene is checked against the oracle (a separate pe world per scene) from an identical
state: contact sets bit-exact, one coloured step equal to sequential Gauss-Seidel in the
batch's colour order (1e-12, correctly rounded trig on both sides), conflict-free colours.
The launch configuration (workgroup size, grid, steps per launch, hipGraph replay) and
contact-capacity growth must not change any result bit.
"""
import numpy as np
import pytest

import oracle as O
from pe_amd import scenes, PeHipError
from pe_amd.batch import BatchWorld

from test_gpu_parity import (contact_map, pair_ranks, rel_err, dense_scene, poly_scene, box_scene, stack_scene,
                             capsule_scene)

pytestmark = pytest.mark.gpu


def small_scene(n, seed, mu=0.4):
    # n^3 spheres resting on the floor of their own box, small random velocities
    return scenes.drop_scene(n, n, n, r=0.5, spacing=1.0 - 2e-6, jitter=2e-7, box=(n + 1.5, n + 1.5, 20.0), mu=mu,
                             seed=seed, v_amp=0.05, max_iterations=30, z0=0.5 - 1e-6)


def batch_scenes():
    return [small_scene(3, 1), dense_scene(), small_scene(2, 3), poly_scene(), box_scene(), small_scene(4, 5),
            stack_scene(), capsule_scene()]


def oracle_states(scs, steps, cr_trig=False):
    out = []
    for sc in scs:
        ow = O.OracleWorld(gravity=sc.gravity, max_iterations=sc.max_iterations, model=3, cr_trig=cr_trig)
        sc.to_oracle(ow)
        if steps:
            ow.step(sc.dt, steps)
        out.append((ow, ow.get_state()))
    return out


def concat_states(sts):
    return {k: np.concatenate([s[k] for s in sts]) for k in sts[0]}


def test_batch_contact_sets_bitexact():
    scs = batch_scenes()
    ows = oracle_states(scs, 7)
    bw = BatchWorld(scs)
    bw.setState(concat_states([st for _, st in ows]))
    bw.run(1)                      # the contacts of a step are detected on its initial state
    for k, (ow, _) in enumerate(ows):
        g = contact_map(bw.contacts(k))
        ex = contact_map(ow.detect(exhaustive=True))
        assert sum(len(v) for v in ex.values()) > 5, k
        assert g == ex, "scene %d (%s)" % (k, scs[k].name)


@pytest.mark.parametrize("model", [3, 5])
def test_batch_step_equals_gs_in_colour_order(model):
    scs = batch_scenes()
    ows = oracle_states(scs, 10, cr_trig=True)
    for ow, _ in ows:
        ow.L.orc_set_relaxation_model(ow.h, model)
    bw = BatchWorld(scs, exact_trig=True, relaxation_model=model)
    bw.setState(concat_states([st for _, st in ows]))
    bw.run(1)
    g1 = bw.getState()
    for k, (ow, _) in enumerate(ows):
        idx, offs = bw.sweepOrder(k)
        gc = bw.contacts(k)
        assert len(offs) >= 3 and offs[-1] == len(idx) == len(gc["b1"])
        ow.set_sweep_order_ranked(gc["b1"][idx], gc["b2"][idx], pair_ranks(gc)[idx])
        ow.step(scs[k].dt)
        o1 = ow.get_state()
        gk = bw.sceneState(g1, k)
        for key in ("v", "w", "pos", "q"):
            assert rel_err(gk[key], o1[key]) <= 1e-12, (k, scs[k].name, key)


@pytest.mark.parametrize("model", [2, 4, 1, 0])
def test_batch_other_models_equal_gs_in_colour_order(model):
    # the batch kernel runs every relaxation model of the single-world path: the projections
    # (2, 4), the approximate decoupling (1) and the frictionless contact (0)
    scs = batch_scenes()
    ows = oracle_states(scs, 10)
    bw = BatchWorld(scs, relaxation_model=model)
    bw.setState(concat_states([st for _, st in ows]))
    bw.run(1)
    g1 = bw.getState()
    for k, (ow, _) in enumerate(ows):
        idx, _ = bw.sweepOrder(k)
        gc = bw.contacts(k)
        ow.L.orc_set_relaxation_model(ow.h, model)
        ow.set_sweep_order_ranked(gc["b1"][idx], gc["b2"][idx], pair_ranks(gc)[idx])
        ow.step(scs[k].dt)
        o1 = ow.get_state()
        gk = bw.sceneState(g1, k)
        for key in ("v", "w", "pos", "q"):
            assert rel_err(gk[key], o1[key]) <= 1e-12, (k, scs[k].name, key)


def test_batch_colouring_is_conflict_free():
    scs = batch_scenes()
    bw = BatchWorld(scs)
    bw.run(3)
    st = bw.sceneStats()
    for k, sc in enumerate(scs):
        idx, offs = bw.sweepOrder(k)
        gc = bw.contacts(k)
        fixed = sc.arrays()["fixed"]
        assert sorted(idx.tolist()) == list(range(len(gc["b1"])))
        for c in range(len(offs) - 1):
            seg = idx[offs[c]:offs[c + 1]]
            bodies = [int(b) for b in np.concatenate([gc["b1"][seg], gc["b2"][seg]]) if not fixed[b]]
            assert len(bodies) == len(set(bodies)), (k, c)
        assert st["colours"][k] == len(offs) - 1 and st["contacts"][k] == len(gc["b1"])
        assert st["steps"][k] == 3


def run_batch(scs, steps, chunks=1, **tuning):
    bw = BatchWorld(scs)
    if tuning:
        bw.setTuning(**tuning)
    states = []
    for _ in range(chunks):
        bw.run(steps)
        states.append(bw.getState())      # eager work between (graph) launches
    return bw, states


@pytest.mark.parametrize("tuning", [dict(threads=64, blocks_per_cu=1, steps_per_launch=1, use_graph=0),
                                    dict(threads=512, blocks_per_cu=4, steps_per_launch=3, use_graph=0),
                                    dict(threads=128, blocks_per_cu=2, steps_per_launch=2, use_graph=1)])
def test_batch_results_independent_of_launch_configuration(tuning):
    scs = batch_scenes()
    _, ref = run_batch(scs, 4, chunks=3, threads=256, blocks_per_cu=2, steps_per_launch=1, use_graph=0)
    _, got = run_batch(scs, 4, chunks=3, **tuning)
    for a, b in zip(got, ref):
        for key in ("pos", "q", "v", "w"):
            assert np.array_equal(a[key], b[key]), key


def test_batch_capacity_growth_is_transparent():
    scs = batch_scenes()
    ref_bw, ref = run_batch(scs, 3)
    bw = BatchWorld(scs, contacts_per_body=0.01)   # 64 slots + a fraction per body: most scenes overflow
    bw.run(3)
    got = bw.getState()
    assert bw.stats()["capacity_retries"] >= 1
    for key in ("pos", "q", "v", "w"):
        assert np.array_equal(got[key], ref[-1][key]), key
    assert np.array_equal(bw.sceneStats()["steps"], np.full(len(scs), 3))


def test_batch_scene_friction_matches_oracle_materials():
    # Per-scene friction mu_s == the pair friction of a material with csf = mu_s / 2.
    mus = [0.2, 0.5, 0.8]
    scs = [small_scene(3, 7 + k, mu=0.4) for k in range(3)]
    ows = []
    for k, mu in enumerate(mus):
        s2 = small_scene(3, 7 + k, mu=mu)
        ow = O.OracleWorld(gravity=s2.gravity, max_iterations=s2.max_iterations, model=3)
        s2.to_oracle(ow)
        ow.step(s2.dt, 5)
        ows.append(ow)
    bw = BatchWorld(scs, scene_friction=(mus, mus))
    bw.setState(concat_states([ow.get_state() for ow in ows]))
    bw.run(1)
    for k, ow in enumerate(ows):
        gc = bw.contacts(k)
        ex = ow.detect(exhaustive=True)
        assert contact_map(gc) == contact_map(ex)
        assert np.all(gc["mu"] == mus[k])


def test_batch_autotune_keeps_state():
    scs = batch_scenes()
    bw = BatchWorld(scs)
    st0 = bw.getState()
    best = bw.autotune(trial_steps=2)
    st1 = bw.getState()
    for key in ("pos", "q", "v", "w"):
        assert np.array_equal(st0[key], st1[key]), key
    assert best["threads"] in (64, 128, 256, 512)
    assert np.all(bw.sceneStats()["steps"] == 0)
    bw.run(2)
    assert np.all(bw.sceneStats()["steps"] == 2)


def test_batch_errors():
    scs = batch_scenes()
    bw = BatchWorld(scs)
    with pytest.raises(PeHipError, match="Invalid time step size"):
        bw.run(1, dt=0.0)
    with pytest.raises(PeHipError):
        bw.setTuning(threads=96)
    big = scenes.drop_scene(17, 17, 15, r=0.5, spacing=1.05, jitter=0.0, box=(20, 20, 40), mu=0.4, seed=1)
    with pytest.raises(PeHipError, match="4096"):
        BatchWorld([big])
