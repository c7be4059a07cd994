"""The adjoint's per-contact map (pe_amd/csrc/adjoint.hpp) on the CPU: its values and its
dual-number Jacobian against an independent numpy restatement of one contact's NBGS update and
central differences of it.

The map is compiled for the host from the same header the device kernel uses
(tests/native/adjoint_map_host.hip).  Each case is a converged contact: velocities before the
contact's impulse, the impulse the update gives, and the velocities after it, so that E2 = 0 at
the evaluation point.  The restatement follows pe's steps (HCSITS.h:1774-1800 frame and D,
2048-2093 model 0, 2141-2290 model 1, 2395-2520 model 3) with its own choices where pe's search
is discrete: model 3's tangential impulse is the global minimiser of 0.5 x.A x + x.g_to on the
friction circle, located by sampling and refined by root finding on its complex-step derivative.
Inverse inertias are anisotropic, so the tangential block A is too and the minimiser is not the
projection of the unconstrained impulse (for isotropic A they coincide).
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy.optimize import brentq

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "pe_amd", "csrc")


def build_host_map(tmp_path_factory, extra=(), opt="-O2"):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc is not available")
    exe = str(tmp_path_factory.mktemp("adjmap") / "adjoint_map_host")
    cmd = [hipcc, "--offload-arch=gfx950", opt, "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", *extra, "-I", CSRC, "-x", "hip",
           os.path.join(HERE, "native", "adjoint_map_host.hip"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    return exe


def case_line(model, K, z):
    vals = [model, K["rad1"], K["rad2"], K["im1"], K["im2"], *K["I1"].ravel(), *K["I2"].ravel(),
            int(K["plane"]), *K["pn"], K["pd"], *K["xp"], K["erp"], K["dtinv"], *z]
    return " ".join(repr(float(v)) if not isinstance(v, int) else str(v) for v in vals)


@pytest.fixture(scope="module")
def host_map(tmp_path_factory):
    exe = build_host_map(tmp_path_factory)

    def run(model, K, z):
        out = subprocess.run([exe], input=case_line(model, K, z) + "\n", capture_output=True, text=True,
                             timeout=60)
        assert out.returncode == 0, out.stderr
        rows = np.array([[float(x) for x in l.split()] for l in out.stdout.strip().splitlines()])
        return rows[0], rows[1:]          # outputs (15), Jacobian (22 x 15)
    run.exe = exe
    return run


# ----------------------------------------------------------------------------- restatement
def contact_geometry(K, x1, x2):
    """Sphere-sphere (MaxContacts.h:547-568) or sphere-plane (MaxContacts.h:777-792) contact."""
    if K["plane"]:
        n = np.asarray(K["pn"], float)
        dist = n @ x1 - K["rad1"] - K["pd"]
        gpos = x1 - (K["rad1"] + dist) * n
        x2 = np.asarray(K["xp"], float)
    else:
        d = x1 - x2
        dist = np.linalg.norm(d) - K["rad1"] - K["rad2"]
        n = d / np.linalg.norm(d)
        gpos = x2 + n * (K["rad2"] + 0.5 * dist)
    return n, dist, gpos - x1, gpos - x2


def tangent(n):
    """HCSITS.h:1774-1786: n x e for the axis e the normal is least aligned with."""
    if abs(n[0]) < abs(n[1]):
        t = np.array([0.0, n[2], -n[1]]) if abs(n[0]) < abs(n[2]) else np.array([n[1], -n[0], 0.0])
    else:
        t = np.array([-n[2], 0.0, n[0]]) if abs(n[1]) < abs(n[2]) else np.array([n[1], -n[0], 0.0])
    return t / np.linalg.norm(t)


def skew(v):
    return np.array([[0.0, -v[2], v[1]], [v[2], 0.0, -v[0]], [-v[1], v[0], 0.0]])


def circle_minimiser(A, g, radius, planted=None):
    """argmin of 0.5 x.A x + x.g over |x| = radius: the best of 4096 sampled angles, refined by
    bracketing a root of the complex-step derivative."""
    sign = -1.0 if planted == "objective_sign" else 1.0

    def fp(a):
        h = 1e-30
        x = radius * np.array([np.cos(a + 1j * h), np.sin(a + 1j * h)])
        return (0.5 * sign * (x @ A @ x) + x @ g).imag / h

    grid = np.linspace(0.0, 2 * math.pi, 4096, endpoint=False)
    xs = radius * np.vstack([np.cos(grid), np.sin(grid)])
    a0 = grid[int(np.argmin(0.5 * sign * np.einsum("ik,ik->k", xs, A @ xs) + g @ xs))]
    step = 2 * math.pi / 4096
    lo, hi = a0 - 2 * step, a0 + 2 * step
    a = brentq(fp, lo, hi, xtol=1e-16, rtol=1e-15, maxiter=200) if fp(lo) < 0 < fp(hi) else a0
    return radius * np.array([math.cos(a), math.sin(a)])


def local_impulse(model, K, x1, x2, gdot, mu, p_prev, planted=None, info=None):
    """One contact's new impulse (world frame) from the relative velocity gdot with its own impulse
    removed; p_prev is the impulse before the update (model 1 reads its tangential part)."""
    n, dist, r1, r2 = contact_geometry(K, x1, x2)
    t = tangent(n)
    o = np.cross(n, t)
    C = np.column_stack([n, t, o])
    if dist < 0:
        dist = dist * K["erp"]
    I1, I2 = K["I1"], K["I2"]
    Km = -(skew(r1) @ I1 @ skew(r1) + skew(r2) @ I2 @ skew(r2)) + (K["im1"] + K["im2"]) * np.eye(3)
    D = C.T @ Km @ C
    g = C.T @ gdot
    g[0] += dist * K["dtinv"]
    if g[0] >= 0:
        return np.zeros(3)
    if model == 0:
        return n * (-(g[0] / D[0, 0]))
    pc = -np.linalg.solve(D, g)
    if not (pc[1] ** 2 + pc[2] ** 2 > (mu * pc[0]) ** 2 or pc[0] < 0):
        return C @ pc

    def coupling(p):
        return np.cross(I1 @ np.cross(r1, p), r1) + np.cross(I2 @ np.cross(r2, p), r2)

    A = D[1:, 1:]
    if model == 1:
        pt = (t @ p_prev) * t + (o @ p_prev) * o
        c = max(0.0, -(g[0] + n @ coupling(pt)) / D[0, 0])
        g2 = gdot + coupling(n * c)
        x = -np.linalg.solve(A, np.array([t @ g2, o @ g2]))
        if x @ x > (mu * c) ** 2:
            x = x * (mu * c / math.sqrt(x @ x))
        return C @ np.array([c, x[0], x[1]])
    x = pc[1:].copy()
    for _ in range(200):                       # decoupled passes to their fixed point
        c = max(0.0, -(n @ (gdot + coupling(t * x[0] + o * x[1])) + dist * K["dtinv"]) / D[0, 0])
        g2 = gdot + coupling(n * c)
        gto = np.array([t @ g2, o @ g2])
        u = -np.linalg.solve(A, gto)
        radius = (1.0 if planted == "radius_without_mu" else mu) * c
        xn = circle_minimiser(A, gto, radius, planted) if u @ u > radius ** 2 else u
        done = np.abs(xn - x).max() <= 1e-15 * max(np.abs(x).max(), 1e-300)
        x = xn
        if done:
            break
    if info is not None:
        info.update(x=x, u=u, radius=radius)
    return C @ np.array([c, x[0], x[1]])


def contact_outputs(model, K, z, planted=None):
    """(E2, -contrib) of contact_map from z = (x1, x2, v1, w1, v2, w2, p, mu)."""
    x1, x2 = z[0:3], z[3:6]
    v1, w1, v2, w2 = z[6:9], z[9:12], z[12:15], z[15:18]
    p, mu = z[18:21], z[21]
    n, dist, r1, r2 = contact_geometry(K, x1, x2)
    dv1 = v1 - K["im1"] * p
    dw1 = w1 - K["I1"] @ np.cross(r1, p)
    dv2 = v2 + K["im2"] * p
    dw2 = w2 + K["I2"] @ np.cross(r2, p)
    gdot = (dv1 - dv2) + np.cross(dw1, r1) - np.cross(dw2, r2)
    pn = local_impulse(model, K, x1, x2, gdot, mu, p, planted)
    contrib = np.concatenate([-K["im1"] * p, -K["I1"] @ np.cross(r1, p), K["im2"] * p, K["I2"] @ np.cross(r2, p)])
    return np.concatenate([p - pn, contrib])


def spd(rng, lo, hi):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q @ np.diag(rng.uniform(lo, hi, 3)) @ q.T


def converged_case(model, seed, plane, sliding=True):
    """A contact at its converged impulse: velocities U = U0 + the impulse's corrections, with the
    impulse the update gives from U0."""
    rng = np.random.default_rng(seed)
    K = {"rad1": rng.uniform(0.3, 0.6), "rad2": rng.uniform(0.3, 0.6), "im1": rng.uniform(0.5, 2.0),
         "im2": 0.0 if plane else rng.uniform(0.5, 2.0), "I1": spd(rng, 2.0, 12.0),
         "I2": np.zeros((3, 3)) if plane else spd(rng, 2.0, 12.0), "plane": plane, "erp": 0.7, "dtinv": 1000.0}
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    if plane:
        K["pn"], K["xp"] = n, np.array([0.1, -0.2, 0.3])
        K["pd"] = float(n @ K["xp"])
        x2 = K["xp"].copy()
        x1 = K["xp"] + (K["rad1"] - 0.0005) * n + 0.3 * np.cross(n, tangent(n))
    else:
        K["pn"], K["xp"], K["pd"] = np.zeros(3), np.zeros(3), 0.0
        x2 = rng.normal(size=3)
        x1 = x2 + (K["rad1"] + K["rad2"] - 0.0005) * n
    t = tangent(n)
    vt = (4.0 if sliding else 0.02) * (np.cos(1.1 * seed) * t + np.sin(1.1 * seed) * np.cross(n, t))
    v1, v2 = -0.3 * n + vt + 0.05 * rng.normal(size=3), np.zeros(3) if plane else 0.05 * rng.normal(size=3)
    w1, w2 = 0.5 * rng.normal(size=3), np.zeros(3) if plane else 0.5 * rng.normal(size=3)
    mu = rng.uniform(0.2, 0.6)
    z0 = np.concatenate([x1, x2, v1, w1, v2, w2, np.zeros(3), [mu]])
    if model == 1:
        # model 1 reads the previous impulse's tangential part: iterate the update to its fixed point
        p = np.zeros(3)
        for _ in range(200):
            z0[18:21] = p
            pn = p - contact_outputs(model, K, z0)[:3]
            p, same = pn, np.array_equal(pn, p)
            if same:
                break
        z0[18:21] = 0.0
    n_, dist, r1, r2 = contact_geometry(K, x1, x2)
    gdot = (v1 - v2) + np.cross(w1, r1) - np.cross(w2, r2)
    p = local_impulse(model, K, x1, x2, gdot, mu, p if model == 1 else np.zeros(3))
    z = z0.copy()
    z[6:9] = v1 + K["im1"] * p
    z[9:12] = w1 + K["I1"] @ np.cross(r1, p)
    z[12:15] = v2 - K["im2"] * p
    z[15:18] = w2 - K["I2"] @ np.cross(r2, p)
    z[18:21] = p
    return K, z


def central_jacobian(model, K, z, h=1e-6):
    J = np.zeros((22, 15))
    for j in range(22):
        if K["plane"] and 3 <= j < 6:
            continue                                    # a plane's position is a constant of the map
        zp, zm = z.copy(), z.copy()
        zp[j] += h
        zm[j] -= h
        J[j] = (contact_outputs(model, K, zp) - contact_outputs(model, K, zm)) / (2 * h)
    return J


CASES = [(3, s, pl) for s in (1, 2, 3, 4) for pl in (0, 1)] + [(1, 5, 0), (1, 6, 1), (0, 7, 0), (3, 8, 0)]


@pytest.mark.parametrize("model,seed,plane", CASES)
def test_map_values_and_jacobian(host_map, model, seed, plane):
    K, z = converged_case(model, seed, plane, sliding=seed != 8)
    out, J = host_map(model, K, z)
    ref = contact_outputs(model, K, z)
    p = z[18:21]
    if model == 3 and seed != 8:
        # a sliding contact: on the circle, and not the projection of the unconstrained impulse
        n, _, r1, r2 = contact_geometry(K, z[0:3], z[3:6])
        pnrm = n @ p
        assert pnrm > 0 and abs(np.linalg.norm(p - pnrm * n) - z[21] * pnrm) < 1e-12 * pnrm
        info = {}
        gdot = (z[6:9] - K["im1"] * p - z[12:15] - K["im2"] * p) + np.cross(z[9:12] - K["I1"] @ np.cross(r1, p), r1) \
            - np.cross(z[15:18] + K["I2"] @ np.cross(r2, p), r2)
        local_impulse(3, K, z[0:3], z[3:6], gdot, z[21], p, info=info)
        proj = info["u"] * (info["radius"] / np.linalg.norm(info["u"]))
        assert np.linalg.norm(info["x"] - proj) > 1e-3 * info["radius"]
    # the converged point: E2 = 0 in both, i.e. the host map's minimiser is the restatement's
    assert np.abs(ref[:3]).max() <= 1e-14 * max(1.0, np.abs(p).max())
    assert np.abs(out[:3]).max() <= 1e-12 * max(1.0, np.abs(p).max()), out[:3]
    assert np.allclose(out[3:], ref[3:], rtol=1e-14, atol=1e-15)
    Jfd = central_jacobian(model, K, z)
    if K["plane"]:
        J = J.copy()
        J[3:6] = 0.0
    scale = np.abs(Jfd).max()
    err = np.abs(J - Jfd).max()
    assert err <= 2e-7 * scale, (err, scale, np.unravel_index(np.argmax(np.abs(J - Jfd)), J.shape))


@pytest.mark.parametrize("planted", ["objective_sign", "radius_without_mu"])
def test_planted_mistakes_are_caught(host_map, planted):
    # a restatement with a planted mistake in model 3's minimiser disagrees with the host map
    K, z = converged_case(3, 2, 0)
    out, _ = host_map(3, K, z)
    bad = contact_outputs(3, K, z, planted=planted)
    assert np.abs(out[:3] - bad[:3]).max() > 1e-6 * np.abs(z[18:21]).max()


def test_map_clean_under_sanitizers(tmp_path_factory, host_map):
    # The same driver built with AddressSanitizer and UndefinedBehaviorSanitizer on its host code
    # (no recovery, leak detection on) runs every case above without a report and prints the same
    # bits as the plain build: no out-of-bounds or freed-memory access, no leak, and none of the
    # undefined behaviour UBSan checks (signed overflow, bad shifts, null or misaligned accesses).
    san = build_host_map(tmp_path_factory, opt="-O1",
                         extra=("-g", "-Xarch_host", "-fsanitize=address,undefined",
                                "-Xarch_host", "-fno-sanitize-recover=undefined"))
    lines = "\n".join(case_line(m, *converged_case(m, s_, pl, sliding=s_ != 8)) for m, s_, pl in CASES) + "\n"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:exitcode=99",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    got = subprocess.run([san], input=lines, capture_output=True, text=True, timeout=300, env=env)
    assert got.returncode == 0 and "runtime error" not in got.stderr and "Sanitizer" not in got.stderr, got.stderr
    plain = subprocess.run([host_map.exe], input=lines, capture_output=True, text=True, timeout=60)
    assert plain.returncode == 0
    assert got.stdout == plain.stdout
    assert len(got.stdout.splitlines()) == len(CASES) * 23
