"""Pins of the oracle's body caching for non-spherical bodies (HCSITS.h:1834-1878, 4094-4109):
world-frame inertia R I R^T (RigidBody.h:842-869) and the gyroscopic predictor
w + dt Iw^-1 ((Iw w) x w) (HCSITS.h:1852).

The expected values come from the mechanics, computed here independently of the oracle's own
formulas: Rodrigues rotation matrices, the box's principal moments, Euler's equations of the
torque-free rigid body (closed-form precession of a symmetric top, conserved world angular
momentum).  Each check is also run against oracle builds with a planted mistake (gyroscopic
cross product reversed, inertia rotated as R^T I R, gyroscopic term left out) to show that the
check fails for them.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

RHO = 2500.0


def rodrigues(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def quat_matrix(q):
    """Rotation matrix of a unit quaternion (w, x, y, z), written out from the Hamilton product."""
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def box_moments(lengths):
    lx, ly, lz = lengths
    m = RHO * lx * ly * lz
    return m, np.array([m / 12 * (ly * ly + lz * lz), m / 12 * (lx * lx + lz * lz), m / 12 * (lx * lx + ly * ly)])


def box_world(lengths, q, w, lib_path=None):
    ow = O.OracleWorld(gravity=(0, 0, 0), max_iterations=10, model=3, lib_path=lib_path)
    mat = ow.create_material(RHO, 0.0, 0.2, 0.2)
    ow.create_box([0, 0, 50.0], lengths, mat)
    st = ow.get_state()
    st["q"][0] = q
    st["w"][0] = w
    ow.set_state(st)
    return ow


# --- checks, each returning its error (so that planted builds can be shown to fail) ----------

def torque_response_error(lib_path=None):
    """One step from rest with a torque: w = dt Iw^-1 tau (HCSITS.h:4094-4109), Iw^-1 = R I^-1 R^T."""
    lengths = (0.9, 0.5, 0.3)
    axis, ang = np.array([0.3, -0.7, 0.5]), 0.83
    q = np.concatenate([[np.cos(ang / 2)], np.sin(ang / 2) * axis / np.linalg.norm(axis)])
    ow = box_world(lengths, q, np.zeros(3), lib_path)
    tau = np.array([[12.0, -30.0, 7.0]])
    ow.set_forces(np.zeros((1, 3)), tau)
    dt = 1e-3
    ow.step(dt)
    w = ow.get_state()["w"][0]
    _, I = box_moments(lengths)
    R = rodrigues(axis, ang)
    expect = dt * (R @ np.diag(1.0 / I) @ R.T @ tau[0])
    return float(np.max(np.abs(w - expect)) / np.max(np.abs(expect)))


def momentum_drift(lib_path=None, steps=4000, dt=2e-4):
    """Torque-free asymmetric box: world angular momentum R I R^T w is conserved by the exact
    dynamics; the explicit gyroscopic step drifts by O(dt)."""
    lengths = (1.0, 0.6, 0.3)
    axis, ang = np.array([1.0, 2.0, -0.5]), 1.1
    q = np.concatenate([[np.cos(ang / 2)], np.sin(ang / 2) * axis / np.linalg.norm(axis)])
    w0 = np.array([0.4, 2.5, -1.0])      # mostly about the intermediate body axis (rotated): tumbling
    ow = box_world(lengths, q, w0, lib_path)
    _, I = box_moments(lengths)

    def L(st):
        R = quat_matrix(st["q"][0])
        return R @ np.diag(I) @ R.T @ st["w"][0]

    L0 = L(ow.get_state())
    ow.step(dt, steps)
    return float(np.linalg.norm(L(ow.get_state()) - L0) / np.linalg.norm(L0))


def precession_error(lib_path=None, dt=1e-4, steps=5000):
    """Symmetric top (I1 = I2): in the body frame w3 stays constant and (w1, w2) turns at
    Omega = (I3 - I1) / I1 * w3 (Euler's equations).  Returns the error of the turned angle
    (radians) and of w3 (relative)."""
    lengths = (0.8, 0.8, 0.3)
    _, I = box_moments(lengths)
    w0 = np.array([0.6, 0.0, 4.0])
    ow = box_world(lengths, [1.0, 0, 0, 0], w0, lib_path)
    ow.step(dt, steps)
    st = ow.get_state()
    wb = quat_matrix(st["q"][0]).T @ st["w"][0]
    omega = (I[2] - I[0]) / I[0] * w0[2]
    turned = np.arctan2(wb[1], wb[0])
    expect = np.arctan2(np.sin(omega * steps * dt), np.cos(omega * steps * dt))
    dang = abs(np.arctan2(np.sin(turned - expect), np.cos(turned - expect)))
    return float(dang), float(abs(wb[2] - w0[2]) / w0[2]), float(omega * steps * dt)


def test_torque_response_rotated_box():
    assert torque_response_error() < 1e-13


def test_angular_momentum_conserved_asymmetric_box():
    # first order: halving dt halves the drift; both far below a planted mistake's O(1)
    d1 = momentum_drift(steps=4000, dt=2e-4)
    d2 = momentum_drift(steps=8000, dt=1e-4)
    assert d1 < 1e-3             # measured 2.2e-4; planted mistakes drift by 0.47-0.60
    assert 0.3 < d2 / d1 < 0.7


def test_symmetric_top_precession():
    dang, dw3, turned = precession_error()
    assert turned > 1.0          # the body-frame vector turns by more than a radian
    assert dang < 1e-5           # measured 4.6e-8 rad after turning 1.507 rad
    assert dw3 < 1e-9


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    """Oracle builds with one planted mistake each (ORC_PLANT in oracle/pe_oracle.cpp)."""
    d = tmp_path_factory.mktemp("planted")
    out = {}
    src = os.path.join(ROOT, "oracle", "pe_oracle.cpp")
    for k in (1, 2, 3):
        so = str(d / ("liboracle_plant%d.so" % k))
        subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
                        "-DORC_PLANT=%d" % k, "-shared", "-o", so, src, "-lquadmath"], check=True)
        out[k] = so
    return out


@pytest.mark.parametrize("k", [1, 2, 3])
def test_planted_predictor_mistakes_fail(planted, k):
    """Every planted mistake fails at least one of the checks above, by a wide margin."""
    lib = planted[k]
    fails = []
    if torque_response_error(lib) >= 1e-13:
        fails.append("torque")
    if momentum_drift(lib) >= 1e-3:
        fails.append("momentum")
    dang, dw3, _ = precession_error(lib)
    if dang >= 1e-5 or dw3 >= 1e-9:
        fails.append("precession")
    assert fails, "planted mistake %d passes every predictor check" % k
    if k in (1, 3):          # gyroscopic mistakes leave the one-step torque response alone
        assert "momentum" in fails and "precession" in fails
    else:                    # the transposed rotation changes every world-frame inertia
        assert "torque" in fails
