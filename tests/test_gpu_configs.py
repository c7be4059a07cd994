"""The batched-scenes configuration at full size: 4,096 scenes of 50-2,000 spheres (log-uniform),
each in its own 10x10x40 m plane box with its own friction coefficient, 40 NBGS sweeps,
dt = 1e-3, 300 steps, the persistent batch kernel replayed from a hipGraph.

Over the whole run every body stays finite and every scene's penetration bounded; from the
batch's state at step 150 a seeded sample of scenes (the smallest and the largest among them)
is checked against the oracle (a pe world per scene): the contact set bit for bit, and one step
with correctly rounded trig equal to the oracle's Gauss-Seidel in the batch's colour order to
1e-12.
"""
import numpy as np
import pytest

import oracle as O
from pe_amd import scenes
from pe_amd.batch import BatchWorld

from test_gpu_parity import contact_map, pair_ranks, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4_run():
    scs, mus = scenes.config4(4096)
    bw = BatchWorld(scs, scene_friction=(mus, mus))
    bw.setTuning(use_graph=1)
    maxpen, mid = [], None
    for k in range(30):
        bw.run(10)
        st = bw.sceneStats()
        assert np.all(st["steps"] == 10 * (k + 1))
        maxpen.append(st["max_penetration"].copy())
        if k == 14:
            mid = bw.getState()
    end = bw.getState()
    bw.close()
    print("C4: 4096 scenes x 300 steps; largest penetration %.3e" % float(np.max(maxpen)), flush=True)
    return scs, mus, mid, end, np.array(maxpen)


def sample(scs):
    sizes = np.array([sc.n_spheres for sc in scs])
    rng = np.random.default_rng(300)
    pick = set(rng.choice(len(scs), 14, replace=False).tolist()) | {int(sizes.argmin()), int(sizes.argmax())}
    return sorted(pick)


def scene_state(scs, st, k):
    off = np.concatenate([[0], np.cumsum([sc.n_bodies for sc in scs])])
    return {key: v[off[k]:off[k + 1]].copy() for key, v in st.items()}


def test_c4_whole_run_finite_and_penetration_bounded(c4_run):
    scs, _, mid, end, maxpen = c4_run
    for st in (mid, end):
        for key in ("pos", "q", "v", "w"):
            assert np.all(np.isfinite(st[key])), key
    # spheres of radius 0.5 resting in piles: overlaps stay a small fraction of the radius
    assert float(np.max(maxpen)) < 0.05
    # every scene has settled onto its floor by the end: no sphere below the floor or above its start
    for k in sample(scs):
        sk = scene_state(scs, end, k)["pos"][6:]
        assert np.all(sk[:, 2] > 0.5 - 0.05) and np.all(sk[:, 2] < 40.0)


def test_c4_sampled_scenes_contact_sets_bitexact(c4_run):
    scs, mus, mid, _, _ = c4_run
    bw = BatchWorld(scs, scene_friction=(mus, mus))
    bw.setState(mid)
    bw.run(1)                          # contacts of the step detected on the step-150 state
    for k in sample(scs):
        ow = O.OracleWorld(gravity=scs[k].gravity, max_iterations=40, model=3)
        scs[k].to_oracle(ow)
        ow.set_state(scene_state(scs, mid, k))
        ex = contact_map(ow.detect(exhaustive=True))
        hg = contact_map(ow.detect(exhaustive=False))
        g = contact_map(bw.contacts(k))
        assert sum(len(v) for v in ex.values()) > 0, k
        assert g == ex and g == hg, "scene %d (%d spheres)" % (k, scs[k].n_spheres)
        assert all(c[3] == mus[k] for v in g.values() for c in v), "scene friction"
    bw.close()


def test_c4_sampled_scenes_step_equals_gs_in_colour_order(c4_run):
    scs, mus, mid, _, _ = c4_run
    bw = BatchWorld(scs, scene_friction=(mus, mus), exact_trig=True)
    bw.setState(mid)
    bw.run(1)
    g1 = bw.getState()
    for k in sample(scs):
        idx, offs = bw.sweepOrder(k)
        gc = bw.contacts(k)
        assert offs[-1] == len(idx) == len(gc["b1"])
        ow = O.OracleWorld(gravity=scs[k].gravity, max_iterations=40, model=3, cr_trig=True)
        scs[k].to_oracle(ow)
        ow.set_state(scene_state(scs, mid, k))
        ow.set_sweep_order_ranked(gc["b1"][idx], gc["b2"][idx], pair_ranks(gc)[idx])
        ow.step(scs[k].dt)
        o1 = ow.get_state()
        gk = scene_state(scs, g1, k)
        for key in ("v", "w", "pos", "q"):
            assert rel_err(gk[key], o1[key]) <= 1e-12, (k, scs[k].n_spheres, key)
    bw.close()
