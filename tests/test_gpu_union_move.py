"""A union moved or rotated through a state upload carries its member spheres as pe's setters do.

pe's Union::setPosition translates every member by the union's displacement (Union.cpp:251-272,
Sphere::update(dp), Sphere.cpp:1031); Union::setOrientation places every member at
pos + R rpos (Union.cpp:306-322, Sphere::update(dq), Sphere.cpp:1059).  A state upload without
member positions must leave the device's members where those setters would, so that the next
step detects the spheres at their new places: checked against an independent numpy placement
(bit for bit), then one step against the oracle given the same member positions (contact set bit
for bit, the step in pe's order to 1e-12).
"""
import numpy as np
import pytest

import oracle as O
from pe_amd import World, SWEEP_REFERENCE_ORDER
from test_gpu_parity import union_scene, oracle_for, contact_map, rel_err

pytestmark = pytest.mark.gpu


def rot_matrix(q):
    w, x, y, z = q
    return np.array([[1 - 2 * y * y - 2 * z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * x * x - 2 * z * z, 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * x * x - 2 * y * y]])


def place(pos, R, rpos):
    # pos + R rpos, each row of R applied left to right as pe's Matrix3x3 * Vector3
    return np.array([pos[k] + ((R[k, 0] * rpos[0] + R[k, 1] * rpos[1]) + R[k, 2] * rpos[2]) for k in range(3)])


def test_moved_and_rotated_unions_carry_members():
    sc = union_scene()
    ow = oracle_for(sc, iters=20, model=1)
    ow.step(sc.dt, 5)
    st0 = ow.get_state()
    mem = ow.members()
    owners = np.unique(mem["owner"])
    u_move, u_rot, u_both = int(owners[0]), int(owners[5]), int(owners[-1])

    w = World()
    sc.to_world(w)
    w.setMaxIterations(20)
    w.setRelaxationModel(1)
    w.setSweepMode(SWEEP_REFERENCE_ORDER)
    w.setState(st0)                                  # members uploaded as given (st0 has mgpos)

    st1 = {k: np.array(v, copy=True) for k, v in st0.items() if k != "mgpos"}
    st1["pos"][u_move] += [0.0125, -0.0375, 0.0]
    ang = 0.7
    q_rot = np.array([np.cos(ang / 2), 0.0, 0.0, np.sin(ang / 2)])          # about z
    st1["q"][u_rot] = q_rot
    st1["pos"][u_both] += [0.05, 0.0, 0.03125]
    q_both = np.array([np.cos(0.3), np.sin(0.3) * 0.6, 0.0, np.sin(0.3) * 0.8])
    st1["q"][u_both] = q_both
    w.setState(st1)                                  # no member positions: the library moves them
    got = w.getState()["mgpos"]

    expect = np.array(mem["gpos"], copy=True)
    for m in np.flatnonzero(mem["owner"] == u_move):
        expect[m] = mem["gpos"][m] + (st1["pos"][u_move] - st0["pos"][u_move])
    for u, q in ((u_rot, q_rot), (u_both, q_both)):
        R = rot_matrix(q)
        for m in np.flatnonzero(mem["owner"] == u):
            expect[m] = place(st1["pos"][u], R, mem["rpos"][m])
    untouched = ~np.isin(mem["owner"], [u_move, u_rot, u_both])
    assert np.array_equal(got[untouched], mem["gpos"][untouched]), "members of unmoved unions changed"
    assert np.array_equal(got, expect), "members not placed as pe's setPosition / setOrientation place them"
    # rigidity: the rotated members keep their distances to the union's centre
    for u in (u_rot, u_both):
        sel = mem["owner"] == u
        assert np.allclose(np.linalg.norm(got[sel] - st1["pos"][u], axis=1), np.linalg.norm(mem["rpos"][sel], axis=1),
                           rtol=0, atol=1e-14)

    ow.set_state(dict(st1, mgpos=expect))
    ow.step(sc.dt)
    c = ow.contacts()
    o1 = ow.get_state()
    ow2 = oracle_for(sc, iters=20, model=1)
    ow2.set_state(dict(st1, mgpos=expect))
    ex = contact_map(ow2.detect(exhaustive=True))
    w.setContactOrder(c["b1"], c["b2"])
    w.simulationStep(sc.dt)
    assert contact_map(w.contacts()) == ex, "contact set after the move differs from the oracle's"
    g1 = w.getState()
    for k in ("v", "w", "pos", "q"):
        assert rel_err(g1[k], o1[k]) <= 1e-12, k
    assert np.array_equal(g1["mgpos"], o1["mgpos"]) or rel_err(g1["mgpos"], o1["mgpos"]) <= 1e-12
