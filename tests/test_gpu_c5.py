"""The differentiable-step configuration at full size: 37^3 = 50,653 spheres (r = 0.4) dropped
as a jittered block onto a floor, mu = 0.5, dt = 5e-4, 400 steps, checkpoints every 20 steps,
relaxation model 1 run to the floating-point fixed point; loss |COM(T) - x*|^2.

The reverse pass's gradients with respect to initial sphere velocities and the friction
coefficient are compared with central differences (h = 1e-6) of plain forward runs to 1e-5
relative, on seeded coordinates (random ones and ones of spheres touching the floor at the end)
whose perturbed runs keep the active set of every step: the contact pairs, each contact's mode
(separating, static, sliding) and the sign of its distance -- pe scales only negative
distances by the error-reduction parameter (HCSITS.h:1795-1798), a kink that the derivative at
fixed active set does not see.
"""
import hashlib

import numpy as np
import pytest

from pe_amd import World, scenes
from pe_amd.adjoint import DifferentiableRun

pytestmark = pytest.mark.gpu

H = 1e-6
MODEL, MU, ITERS, ABORT = 1, 0.5, 20000, 1e-15
# The adjoint sweeps run to an absolute change of 1e-20 in the adjoint impulses (whose scale here
# is ~1e-5): Gauss-Seidel converges slowly on the landed block (44 of its 400 forward solves stall
# at rounding level short of 1e-15 even after 200,000 sweeps), and stopped at 1e-15 the adjoint
# of a floor-touching sphere's velocity was off by 3.3e-5 against central differences that agree
# between h = 1e-6 and 1e-7 to 4.5e-7; at 1e-19 and below it agrees to 4.6e-7.
ADJ_TOL = 1e-20


def signature(w):
    c = w.contacts()
    p = w.impulses()
    pn = np.einsum("ij,ij->i", p, c["normal"])
    pt = np.linalg.norm(p - pn[:, None] * c["normal"], axis=1)
    mode = np.where(np.all(p == 0, axis=1), 0, np.where(pt < c["mu"] * pn * (1 - 1e-9), 1, 2)).astype(np.uint8)
    mode = mode | (np.asarray(c["dist"] < 0, np.uint8) << 2)
    key = (np.asarray(c["b1"], np.uint64) << np.uint64(32)) | np.asarray(c["b2"], np.uint64)
    i = np.argsort(key, kind="stable")
    return hashlib.sha1(key[i].tobytes() + mode[i].tobytes()).hexdigest()


def plain(sc, v0, mu):
    npl = len(sc.planes)
    w = World()
    sc.to_world(w)
    w.setMaxIterations(ITERS)
    w.setRelaxationModel(MODEL)
    w.setAbortThreshold(ABORT)
    w.setPairFriction(0, 0, mu, mu)
    w.host_arrays()["v"][npl:] = v0
    sig = []
    for _ in range(sc.steps):
        w.run(1, sc.dt)
        sig.append(signature(w))
    pos = w.getState()["pos"][npl:]
    w.close()
    return pos, sig


@pytest.fixture(scope="module")
def c5():
    sc = scenes.config5()
    npl, n = len(sc.planes), sc.n_spheres
    assert n == 50653 and sc.steps == 400
    m = 1.0 / sc.arrays()["inv_mass"][npl:]
    M = m.sum()
    v0 = np.zeros((n, 3))
    base, base_sig = plain(sc, v0, MU)
    target = (m[:, None] * base).sum(0) / M + np.array([0.3, -0.2, 0.1])
    run = DifferentiableRun(sc, sc.steps, checkpoint_every=20, model=MODEL, max_iterations=ITERS,
                            abort_threshold=ABORT, adjoint_sweeps=ITERS, adjoint_tol=ADJ_TOL)
    st = run.forward(v0, MU)
    same_forward = bool(np.array_equal(st["pos"][npl:], base))
    d = (m[:, None] * base).sum(0) / M - target
    gpos = np.zeros((npl + n, 3))
    gpos[npl:] = 2.0 * d[None, :] * (m / M)[:, None]
    _, gv, _, gmu = run.backward(gpos)
    stats = dict(run.last_stats)
    run.close()
    touching = np.flatnonzero(base[:, 2] < 0.4 + 1e-6)
    return {"sc": sc, "npl": npl, "n": n, "m": m, "M": M, "v0": v0, "base": base, "sig": base_sig,
            "target": target, "d": d, "L": float(d @ d), "gv": gv, "gmu": gmu, "stats": stats,
            "same_forward": same_forward, "touching": touching, "rows": []}


def test_c5_reverse_pass(c5):
    # the checkpointed forward is the plain run, every step is differentiated, the adjoint
    # sweeps converged
    assert c5["same_forward"]
    st = c5["stats"]
    assert st["steps"] == 400 and st["max_contacts"] > 2000 and st["max_last_delta"] <= ADJ_TOL
    assert np.all(np.isfinite(c5["gv"])) and np.isfinite(c5["gmu"])


def coords():
    n = 50653
    rng = np.random.default_rng(64)
    out = [("mu", None)] + [(int(f) // 3, int(f) % 3) for f in rng.choice(3 * n, 2, replace=False)]
    return out + [("touching", k) for k in range(3)]


@pytest.mark.parametrize("coord", coords(), ids=lambda c: "%s-%s" % c)
def test_c5_gradient_matches_central_difference(c5, coord):
    sc, npl, m, M, v0, target = c5["sc"], c5["npl"], c5["m"], c5["M"], c5["v0"], c5["target"]
    if coord[0] == "touching":
        rng = np.random.default_rng(640 + coord[1])
        coord = (int(rng.choice(c5["touching"])), int(rng.integers(0, 3)))
    runs, same = [], True
    for sgn in (1, -1):
        if coord[0] == "mu":
            pos, sig = plain(sc, v0, MU + sgn * H)
        else:
            vp = v0.copy()
            vp[coord] += sgn * H
            pos, sig = plain(sc, vp, MU)
        same = same and sig == c5["sig"]
        runs.append(pos)
    if not same:
        pytest.skip("the perturbed runs change the active set: central differences not comparable")
    # L+ - L- = (d+ - d-) . (d+ + d-), d+ - d- summed from per-sphere differences (most spheres end
    # identically in both runs, so there is no cancellation)
    dp = (m[:, None] * (runs[0] - runs[1])).sum(0) / M
    ds = ((m[:, None] * runs[0]).sum(0) + (m[:, None] * runs[1]).sum(0)) / M - 2 * target
    g_fd = float(dp @ ds) / (2 * H)
    g = c5["gmu"] if coord[0] == "mu" else float(c5["gv"][npl + coord[0], coord[1]])
    # relative to the larger of the two, floored at what a single free sphere's coordinate gives,
    # 2 |d| T / n (mu barely moves the centre of mass here: floored at the loss)
    s_free = 2.0 * float(np.linalg.norm(c5["d"])) * sc.steps * sc.dt / c5["n"]
    scale = max(abs(g_fd), abs(g), c5["L"] if coord[0] == "mu" else s_free)
    rel = abs(g - g_fd) / scale
    c5["rows"].append((coord, rel))
    assert rel <= 1e-5, (coord, g, g_fd, rel)


def test_c5_enough_coordinates_compared(c5):
    compared = [c for c, _ in c5["rows"]]
    touching = set(c5["touching"].tolist())
    assert len(compared) >= 4 and any(c[0] != "mu" and c[0] in touching for c in compared), compared
