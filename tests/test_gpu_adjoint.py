"""Adjoint of the time step against central finite differences (GPU).

A small block of spheres with random initial velocities lands on a floor; the loss is the
squared distance of the final centre of mass from a target.  Gradients with respect to initial
sphere velocities and the friction coefficient come from the reverse pass (implicit
differentiation of the converged impulses at fixed active set) and are compared with central
differences (h = 1e-6) of the forward run to 1e-5 relative, on coordinates whose perturbed runs
keep every step's contact set and every contact's mode (separating / static / sliding: the
active set) unchanged.
"""
import numpy as np
import pytest
import torch

from pe_amd import World, scenes, PeHipError
from pe_amd.adjoint import DifferentiableRun, FinalPositions

pytestmark = pytest.mark.gpu

STEPS = 60
H = 1e-6
ITERS = 20000    # the solve runs to the floating-point fixed point: finite differences of a run
ABORT = 1e-17    # whose sweeps stop early would carry the truncation error (/ 2h)
# Model 3 is differentiated at the exact minimiser on the friction circle.  Its forward search
# compares objective values, which resolves the angle only to ~sqrt(eps) (the objective is flat
# at the minimum): central differences of model-3 runs at h = 1e-6 are off by up to 57% (measured),
# at any number of search steps.  For spheres the circle's quadratic form is isotropic, so the
# minimiser is the unconstrained impulse projected onto the circle, which is model 1's update:
# model 3's exact fixed point is model 1's, and its adjoint is checked against differences of
# model-1 runs.  80 search steps put the model-3 forward run within ~1e-8 of that point.
MAX_SUB = {0: 20, 1: 20, 3: 80}
ABORTS = {0: ABORT, 1: ABORT, 3: 1e-15}
FD_MODEL = {0: 0, 1: 1, 3: 1}


def block_scene(model=1):
    # 3x3x3 spheres r = 0.4, spacing 0.79 (touching neighbours), the bottom layer 1 mm above the
    # floor, random initial velocities; mu = 0.5
    rng = np.random.default_rng(5)
    pos = scenes.lattice(3, 3, 3, 0.79, (1.0, 1.0, 0.401), 0.002, rng)
    n = len(pos)
    sc = scenes.Scene("adjoint block", [{"name": "g", "density": 2500.0, "cor": 0.0, "csf": 0.25, "cdf": 0.25}],
                      [((0, 0, 1), 0.0, 0)], pos, np.full(n, 0.4), np.zeros(n, np.uint32), dt=5e-4, steps=STEPS,
                      max_iterations=4000)
    sc.model = model
    v0 = rng.uniform(-0.3, 0.3, size=(n, 3))
    return sc, v0


def loss_and_grad_pos(state, sc, target):
    a = sc.arrays()
    m = np.where(a["inv_mass"] > 0, 1.0 / np.where(a["inv_mass"] > 0, a["inv_mass"], 1.0), 0.0)
    M = m.sum()
    com = (m[:, None] * state["pos"]).sum(0) / M
    d = com - target
    return float(d @ d), (2.0 * d)[None, :] * (m / M)[:, None]


def run_plain(sc, v0, mu, model):
    """Forward run step by step with a World (the same steps as the differentiable forward);
    returns the final state and the per-step contact sets."""
    w = World()
    sc.to_world(w)
    w.setMaxIterations(ITERS)
    w.setRelaxationModel(model)
    w.setAbortThreshold(ABORTS[model])
    w.setMaxSubIterations(MAX_SUB[model])
    w.setPairFriction(0, 0, mu, mu)
    a = w.host_arrays()
    a["v"][1:] = v0
    sets = []
    for _ in range(STEPS):
        w.run(1, sc.dt)
        c = w.contacts()
        p = w.impulses()
        pn = np.einsum("ij,ij->i", p, c["normal"])
        pt = np.linalg.norm(p - pn[:, None] * c["normal"], axis=1)
        # active mode of every contact: separating, static (inside the cone), sliding (on it)
        mode = np.where(np.all(p == 0, axis=1), 0, np.where(pt < c["mu"] * pn * (1 - 1e-9), 1, 2))
        sets.append(frozenset(zip(c["b1"].tolist(), c["b2"].tolist(), mode.tolist())))
    st = w.getState()
    w.close()
    return st, sets


@pytest.mark.parametrize("model", [1, 0, 3])
def test_adjoint_matches_central_differences(model):
    sc, v0 = block_scene(model)
    mu = 0.5
    fdm = FD_MODEL[model]
    base, base_sets = run_plain(sc, v0, mu, fdm)
    target = base["pos"][1:].mean(0) + np.array([0.3, -0.2, 0.1])
    run = DifferentiableRun(sc, STEPS, checkpoint_every=16, model=model, max_iterations=ITERS,
                            abort_threshold=ABORTS[model], adjoint_sweeps=ITERS, adjoint_tol=1e-17,
                            max_sub_iterations=MAX_SUB[model])
    st = run.forward(v0, mu)
    if fdm == model:
        assert np.array_equal(st["pos"], base["pos"])        # the checkpointed forward is the plain run
    else:
        own, own_sets = run_plain(sc, v0, mu, model)
        assert np.array_equal(st["pos"], own["pos"]) and own_sets == base_sets
        gap = np.abs(st["pos"] - base["pos"]).max()
        print("model 3 (80 search steps) against model 1: max position difference %.3e" % gap)
        assert gap < 1e-7                                    # model 3 meets model 1's fixed point
    L, gpos = loss_and_grad_pos(st, sc, target)
    gx, gv, _, gmu = run.backward(gpos)
    assert sum(len(s) for s in base_sets) > 100
    sliding = sum(1 for s in base_sets for (_, _, m) in s if m == 2)
    print("contacts over the run", sum(len(s) for s in base_sets), "sliding", sliding)
    if model != 0:
        assert sliding > 0                  # the sliding branch is differentiated too
    rng = np.random.default_rng(11)
    coords = [(int(i), int(k)) for i, k in zip(rng.integers(0, len(v0), 8), rng.integers(0, 3, 8))]
    rows = []
    for (i, k) in coords + [("mu", None)]:
        fd = []
        same = True
        for sgn in (1, -1):
            if i == "mu":
                s, sets = run_plain(sc, v0, mu + sgn * H, fdm)
            else:
                vp = v0.copy()
                vp[i, k] += sgn * H
                s, sets = run_plain(sc, vp, mu, fdm)
            same = same and sets == base_sets
            fd.append(loss_and_grad_pos(s, sc, target)[0])
        if not same:
            continue                        # a contact or a contact's mode changed: FD not comparable
        g_fd = (fd[0] - fd[1]) / (2 * H)
        g = gmu if i == "mu" else gv[1 + i, k]
        if i == "mu" and model == 0:
            # frictionless contacts never read mu: both derivatives are exactly zero
            assert g == 0.0 and g_fd == 0.0, (g, g_fd)
            continue
        # relative to the coordinate's value or, for a coordinate much smaller than the others,
        # to the gradient's max-norm: central differences with h = 1e-6 of a loss of size ~0.1
        # carry ~1e-10 of rounding noise, above 1e-5 of the smallest components
        scale = abs(gmu) if i == "mu" else max(abs(g_fd), np.abs(gv[1:]).max())
        rows.append((i, k, g, g_fd, abs(g - g_fd) / scale))
    for (i, k) in coords[:4]:                                 # initial positions
        fd = []
        for sgn in (1, -1):
            sc2, _ = block_scene(model)
            sc2.sphere_pos = sc2.sphere_pos.copy()
            sc2.sphere_pos[i, k] += sgn * H
            s, sets = run_plain(sc2, v0, mu, fdm)
            fd.append(loss_and_grad_pos(s, sc, target)[0] if sets == base_sets else np.nan)
        if np.isnan(fd).any():
            continue
        g_fd = (fd[0] - fd[1]) / (2 * H)
        rows.append(("x%d" % i, k, gx[1 + i, k], g_fd, abs(gx[1 + i, k] - g_fd) / max(abs(g_fd), np.abs(gx[1:]).max())))
    print("\n".join("%s %s adjoint %.12e fd %.12e rel %.2e" % r for r in rows))
    print("adjoint stats", run.last_stats)
    assert len(rows) >= 6
    assert max(r[4] for r in rows) <= 1e-5
    run.close()


def test_autograd_function_matches_reverse_pass():
    sc, v0 = block_scene(1)
    run = DifferentiableRun(sc, 30, checkpoint_every=8, model=1)
    v = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    pos = FinalPositions.apply(v, mu, run)
    target = torch.tensor([1.5, 1.2, 0.8], dtype=torch.float64)
    m = torch.ones(len(v0), dtype=torch.float64)
    com = (pos[1:] * m[:, None]).sum(0) / m.sum()
    loss = ((com - target) ** 2).sum()
    loss.backward()
    gpos = np.zeros((len(v0) + 1, 3))
    gpos[1:] = (2.0 * (com.detach().numpy() - target.numpy()) / len(v0))[None, :]
    _, gv, _, gmu = run.backward(gpos)
    assert np.allclose(v.grad.numpy(), gv[1:], rtol=1e-12, atol=0)
    assert abs(mu.grad.item() - gmu) <= 1e-12 * max(abs(gmu), 1e-30)
    run.close()


def test_adjoint_rejects_unsupported_models():
    sc, v0 = block_scene(2)
    run = DifferentiableRun(sc, 5, model=2)
    with pytest.raises(PeHipError, match="models 0, 1 and 3"):
        run.forward(v0, 0.5)
    run.close()
