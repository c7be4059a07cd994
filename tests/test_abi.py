"""CPU checks of the C-ABI library: it loads, exports every function declared in
include/pe_hip.h, and its host-only entry points behave (no kernel launches here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from pe_amd import _abi
from pe_amd.world import World, mat3_inverse

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "pe_hip.h")).read()
    return sorted(set(re.findall(r"\b(pehip_[a-z_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(_abi.EXPORTED)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(_abi.LIB_PATH)
    for name in header_functions():
        assert hasattr(L, name), name


def test_default_params_match_pe():
    L = _abi.load()
    p = _abi.Params()
    L.pehip_default_params(ctypes.byref(p))
    # HCSITS.h:408-431 and Settings.cpp:46-48
    assert p.max_iterations == 100 and p.erp == 0.7 and p.max_sub_iterations == 20
    assert p.relaxation_model == 1 and p.relaxation_param == 0.9 and p.damping == 1.0
    assert p.abort_threshold == 0.0 and p.sweep_mode == 0
    assert L.pehip_version().decode().startswith("pe_hip")


def test_invalid_create_arguments():
    L = _abi.load()
    p = _abi.Params()
    L.pehip_default_params(ctypes.byref(p))
    p.relaxation_model = 5
    h = ctypes.c_void_p()
    st = L.pehip_create(ctypes.byref(p), 0, ctypes.byref(h))
    assert st in (_abi.NOT_SUPPORTED, _abi.HIP_ERROR)
    assert not h.value
    assert len(L.pehip_last_error(None)) > 0


def test_world_host_mass_properties_match_oracle(oracle_mod):
    # The host-side factories compute the same bits as the oracle's restatement of pe.
    from pe_amd import scenes
    sc = scenes.drop_scene(4, 4, 3, r=0.37)
    ow = oracle_mod.OracleWorld()
    sc.to_oracle(ow)
    info = ow.body_info()
    a = sc.arrays()
    assert np.array_equal(a["inv_mass"], info["inv_mass"])
    assert np.array_equal(a["Iinv"], info["inv_inertia"])
    st = ow.get_state()
    assert np.array_equal(a["q"], st["q"]) and np.array_equal(a["normal"], st["normal"])
    assert np.array_equal(a["pos"], st["pos"]) and np.array_equal(a["d"], st["d"])


def test_world_api_errors():
    w = World()
    with pytest.raises(ValueError):
        w.createMaterial("m", -1.0, 0, 0.1, 0.1)
    m = w.createMaterial("m", 1000.0, 0, 0.1, 0.1)
    with pytest.raises(ValueError):
        w.createMaterial("m", 1000.0, 0, 0.1, 0.1)
    with pytest.raises(ValueError):
        w.createSphere(0, (0, 0, 0), -1.0, m)
    with pytest.raises(ValueError):
        w.createPlane(0, (0, 0, 0), 1.0, m)
    with pytest.raises(ValueError):
        w.run(1, 0.0)
    assert w._csf[0, 0] == 0.2


def test_world_refuses_arrays_of_the_wrong_size_before_the_library():
    # The C-ABI reads exactly 3n (4n for q, n for d) values per state array, 3n per force array
    # and as many second bodies as first bodies: a shorter array would be read past its end, so
    # the host refuses it (before touching a device, so no GPU is needed here).
    w = World()
    m = w.createMaterial("m", 1000.0, 0, 0.1, 0.1)
    for k in range(3):
        w.createSphere(k, (2.0 * k, 0, 1), 0.5, m)
    w.createPlane(9, (0, 0, 1), 0.0, m)
    n = w.size
    good = {"pos": np.zeros((n, 3)), "q": np.tile([1.0, 0, 0, 0], (n, 1)), "v": np.zeros((n, 3)),
            "w": np.zeros((n, 3)), "normal": np.zeros((n, 3)), "d": np.zeros(n)}
    for key in good:
        bad = dict(good)
        bad[key] = good[key][:-1]
        with pytest.raises(ValueError, match=key):
            w.setState(bad)
    with pytest.raises(ValueError, match="force"):
        w.setForces(np.zeros((n - 1, 3)), np.zeros((n, 3)))
    with pytest.raises(ValueError, match="torque"):
        w.setForces(np.zeros((n, 3)), np.zeros(3 * n + 3))
    with pytest.raises(ValueError, match="contact order"):
        w.setContactOrder([0, 1], [3])
    assert w._ctx is None   # nothing reached the library
    # bulk creation: every body array holds its fixed count per body
    from pe_amd import scenes
    a = scenes.drop_scene(2, 2, 1, r=0.5, spacing=1.2, jitter=0.0, box=(4, 4, 8), seed=1).arrays()
    for key in ("pos", "q", "I", "d"):
        bad = dict(a)
        bad[key] = a[key][:-1]
        with pytest.raises(ValueError, match=key):
            World().add_bodies(bad)
    assert World().add_bodies(a).size == len(a["shape"])


def test_mat3_inverse_host():
    I = np.zeros(9)
    I[0], I[4], I[8] = 2.0, 3.0, 5.0
    inv = mat3_inverse(I)
    assert np.allclose(inv[[0, 4, 8]], [0.5, 1 / 3, 0.2], rtol=1e-15)


def test_ctypes_structs_match_header_layout(tmp_path):
    # The ctypes mirrors in pe_amd/_abi.py have the size and field offsets that a C compiler
    # gives the structs of include/pe_hip.h (gcc on the header alone; no GPU, no library).
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    mirrors = {"pehip_params": _abi.Params, "pehip_stats": _abi.Stats, "pehip_batch_tuning": _abi.BatchTuning,
               "pehip_batch_stats": _abi.BatchStats, "pehip_adjoint_stats": _abi.AdjointStats}
    lines = ['#include <stdio.h>', '#include "pe_hip.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append('  printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('  printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    for cname, py in mirrors.items():
        assert int(got[cname + " size"]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got["%s.%s" % (cname, f)]) == getattr(py, f).offset, (cname, f)


def test_cpp_dropin_host_arrays_equal_python_world(tmp_path):
    # The C++ drop-in (include/pe_amd/pe.hpp) and the Python World build the arrays the library
    # receives on the host: the same factory calls (planes, a sphere, a fixed box, a capsule, a
    # three-sphere union of two materials, a one-sphere union) must give the same bits in every
    # body and member array, system IDs included; the drop-in's union setters then move the
    # members rigidly, and destroying bodies leaves pe's body order on both.  Host only: nothing is
    # stepped.
    import shutil
    import subprocess
    if shutil.which("g++") is None or not os.path.exists(_abi.LIB_PATH):
        pytest.skip("g++ or the library not available")
    exe = str(tmp_path / "dropin_host_arrays")
    lib = os.path.dirname(_abi.LIB_PATH)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "dropin_host_arrays.cpp"), "-o", exe, "-L", lib, "-lpehip",
                    "-Wl,-rpath," + lib], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True, timeout=60).stdout
    cpp = {l.split()[0]: np.array([float(x) for x in l.split()[1:]]) for l in out.splitlines()}

    w = World()
    gran = w.createMaterial("granular", 2500, 0, 0.2, 0.2)
    wood = w.createMaterial("wood", 700, 0.1, 0.35, 0.3)
    w.createPlane(100, (0, 0, 1), 0, gran)
    w.createPlane(101, (1, -2, 0.5), -3.25, wood)
    w.createSphere(1, (1.5, -0.25, 2), 0.375, gran)
    box = w.createBox(2, (-1, 2, 0.75), (0.5, 1.25, 0.8), wood)
    w.createCapsule(3, (0.3, 0.1, 4), 0.2, 1.1, gran)
    w.createUnion(4, (2, 2, 3), [((0.3, 0, 0), 0.25, gran), ((-0.1, 0.2, 0.05), 0.3, wood),
                                 ((0, -0.15, 0.2), 0.2, gran)])
    w.createUnion(5, (-2, -1, 1.5), [((0.1, 0.1, 0.1), 0.4, wood)])
    w.createSphere(6, (0, 0, 7), 0.5, wood)
    w.setFixed(box)
    a, m = w.host_arrays(), w.members()
    py = {k: a[k] for k in ("shape", "sid", "mat", "fixed", "param", "pos", "q", "inv_mass", "I", "Iinv", "normal",
                            "d")}
    py.update({"m_owner": m["owner"], "m_sid": m["sid"], "m_mat": m["mat"], "m_rad": m["radius"],
               "m_rpos": m["rpos"], "m_gpos": m["gpos"]})
    assert set(py) <= set(cpp)
    for k in py:
        assert np.array_equal(cpp[k], np.asarray(py[k], float).ravel()), k
    assert int(cpp["m_owner"][-1]) == 6 and len(cpp["m_sid"]) == 4   # the unions' members follow them
    # World::Iterator and World::Bodies::CastIterator<T>: body order, one type at a time
    assert cpp["iter_all"].tolist() == [100, 101, 1, 2, 3, 4, 5, 6]
    assert cpp["iter_box"].tolist() == [2] and cpp["iter_sphere"].tolist() == [1, 6]
    assert cpp["iter_capsule"].tolist() == [3] and cpp["iter_plane"].tolist() == [100, 101]
    assert cpp["iter_union"].tolist() == [4, 5] and cpp["iter_union_size"].tolist() == [3, 1]

    # Union::setPosition translates the members by the union's displacement; setOrientation puts
    # each at pos + R(q) rpos (Union.cpp:251-272, 306-322).  Checked against a rigid motion built
    # independently: the displacement, and the textbook rotation matrix of the unit quaternion.
    g0, rp = cpp["m_gpos"].reshape(-1, 3), cpp["m_rpos"].reshape(-1, 3)
    mine = cpp["m_owner"] == 5
    moved = cpp["moved_m_gpos"].reshape(-1, 3)
    dp = np.array([2.5, 1.75, 3.25]) - np.array([2.0, 2.0, 3.0])
    assert np.allclose(moved[mine], g0[mine] + dp, rtol=0, atol=1e-15)
    assert np.array_equal(moved[~mine], g0[~mine])
    qw, qx, qy, qz = cpp["turned_q"].reshape(-1, 4)[5]
    assert abs(qw * qw + qx * qx + qy * qy + qz * qz - 1) < 1e-15 and abs(qw - np.cos(0.35)) < 1e-15
    R = np.array([[1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy)],
                  [2 * (qx * qy + qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx)],
                  [2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), 1 - 2 * (qx * qx + qy * qy)]])
    x = cpp["turned_pos"].reshape(-1, 3)[5]
    turned = cpp["turned_m_gpos"].reshape(-1, 3)
    assert np.allclose(turned[mine], x + rp[mine] @ R.T, rtol=0, atol=1e-14)
    assert np.array_equal(turned[~mine], g0[~mine])
    # the members keep their distances to each other (a rigid motion)
    d0 = np.linalg.norm(g0[mine][:, None] - g0[mine][None], axis=-1)
    d1 = np.linalg.norm(turned[mine][:, None] - turned[mine][None], axis=-1)
    assert np.allclose(d0, d1, rtol=0, atol=1e-14)

    # destroy (RigidBody.cpp:766-775): pe's BodyStorage::remove (BodyStorage.h:563-573) moves the
    # last body into the removed one's place.  Removing body 2 (a sphere) then body 5 (the
    # three-sphere union) leaves the bodies [0, 1, 7, 3, 4, 6] of the original order, and the
    # one-sphere union's member now belongs to body 5.
    w.destroy(2)
    w.destroy(5)
    perm = [0, 1, 7, 3, 4, 6]
    a2, m2 = w.host_arrays(), w.members()
    for k in py:
        if not k.startswith("m_"):
            assert np.array_equal(np.asarray(a2[k]), np.asarray(py[k])[perm]), k
            assert np.array_equal(cpp["after_" + k], np.asarray(a2[k], float).ravel()), k
    assert m2["owner"].tolist() == [5] and m2["sid"].tolist() == [int(py["m_sid"][-1])]
    assert np.array_equal(m2["gpos"], np.asarray(py["m_gpos"]).reshape(-1, 3)[-1:])
    for k, mk in (("m_owner", "owner"), ("m_sid", "sid"), ("m_mat", "mat"), ("m_rad", "radius"),
                  ("m_rpos", "rpos"), ("m_gpos", "gpos")):
        assert np.array_equal(cpp["after_" + k], np.asarray(m2[mk], float).ravel()), k
    assert w.unionMass(5) > 0 and 6 not in w._u_mass
    # the C++ handles follow their bodies: World::operator[] walks the new order
    assert cpp["after_handle_index"].tolist() == list(range(6))
    assert cpp["after_handle_uid"].tolist() == [100, 101, 6, 2, 3, 5]
    with pytest.raises(IndexError):
        w.destroy(6)


def test_cpp_dropin_pose_utilities_follow_pe(tmp_path, oracle_mod):
    # pe's translate / rotate / rotateAroundOrigin / rotateAroundPoint / setPosition / setOrientation
    # and pointFromBFtoWF in the C++ drop-in (tests/native/dropin_rotations.cpp), replayed here
    # with the oracle's restatements of pe's quaternion product, Quaternion(axis, angle), rotate and
    # toRotationMatrix (pinned by closed forms in test_oracle.py), step by step in pe's order
    # (Box.cpp:456-870, Plane.cpp:216-880, Union.cpp:563-1100, Quaternion.h:309-315, 645-697):
    # every pose, plane normal and displacement and union member position bit for bit.
    import shutil
    import subprocess
    if shutil.which("g++") is None or not os.path.exists(_abi.LIB_PATH):
        pytest.skip("g++ or the library not available")
    O = oracle_mod
    exe = str(tmp_path / "dropin_rotations")
    lib = os.path.dirname(_abi.LIB_PATH)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "dropin_rotations.cpp"), "-o", exe, "-L", lib, "-lpehip",
                    "-Wl,-rpath," + lib], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True, timeout=60).stdout
    got = {l.split()[0]: [float(x) for x in l.split()[1:]] for l in out.splitlines()}
    rpos = np.array(got["m_rpos"]).reshape(-1, 3)

    Q = O.quat_from_axis_angle
    mul = O.quat_mul

    def euler(x, y, z):          # Quaternion(xangle, yangle, zangle): identity turned about x, y, z
        return mul(Q((0, 0, 1), z), mul(Q((0, 1, 0), y), mul(Q((1, 0, 0), x), np.array([1.0, 0, 0, 0]))))

    def turned_about(pos, dq, point):
        dp = [pos[k] - point[k] for k in range(3)]
        r = O.quat_rotate(dq, dp)
        return np.array([point[k] + r[k] for k in range(3)])

    def members(pos, q):
        R = O.quat_to_rotation(q)
        return np.array([[pos[k] + (R[k, 0] * r[0] + R[k, 1] * r[1] + R[k, 2] * r[2]) for k in range(3)] for r in rpos])

    def check(step, pos, q, normal=(0.0, 0.0, 0.0), d=0.0, mem=None):
        row = np.array(got[step])
        assert np.array_equal(row[:3], pos), (step, "pos")
        assert np.array_equal(row[3:7], q), (step, "q")
        assert np.array_equal(row[7:10], normal), (step, "normal")
        assert row[10] == d, (step, "d")
        if mem is not None:
            assert np.array_equal(row[11:].reshape(-1, 3), mem), (step, "members")

    # box
    pos, q = np.array([1.25, -0.5, 2.0]), np.array([1.0, 0, 0, 0])
    check("box0", pos, q)
    q = mul(Q((1, 2, 3), 0.4), q); check("box1", pos, q)
    q = mul(Q((0, 0, 1), 0.5), mul(Q((0, 1, 0), -0.2), mul(Q((1, 0, 0), 0.3), q))); check("box2", pos, q)
    q = mul(Q((-1, 0.5, 2), 1.1), q); check("box3", pos, q)
    dq = Q((0, 1, 1), 0.7); pos = O.quat_rotate(dq, pos); q = mul(dq, q); check("box4", pos, q)
    dq = euler(0.2, 0.1, -0.3); pos = O.quat_rotate(dq, pos); q = mul(dq, q); check("box5", pos, q)
    dq = Q((0.3, 0.3, 1), 0.9); pos = turned_about(pos, dq, (1, -1, 0.5)); q = mul(dq, q); check("box6", pos, q)
    dq = euler(0.4, -0.6, 0.2); pos = turned_about(pos, dq, (-0.5, 0.25, 1)); q = mul(dq, q); check("box7", pos, q)
    pos = np.array([pos[0] + 0.1, pos[1] + 0.2, pos[2] + -0.3]); check("box8", pos, q)
    R = O.quat_to_rotation(q)
    # getRotation is toRotationMatrix of the quaternion; getInertia / getInvInertia are
    # R I trans(R) in pe's product order (RigidBody.h:842-869, Matrix3x3.h:2241-2253)
    assert got["box_R"] == R.ravel().tolist()
    Ib, Ibinv = np.array(got["box_Ib"]).reshape(3, 3), np.array(got["box_Ibinv"]).reshape(3, 3)
    assert Ib[0, 0] > 0 and np.allclose(Ib @ Ibinv, np.eye(3), rtol=0, atol=1e-15)

    def rirt(M):
        t = [[R[i, 0] * M[0, j] + R[i, 1] * M[1, j] + R[i, 2] * M[2, j] for j in range(3)] for i in range(3)]
        return [t[i][0] * R[j, 0] + t[i][1] * R[j, 1] + t[i][2] * R[j, 2] for i in range(3) for j in range(3)]
    assert got["box_Iw"] == rirt(Ib) and got["box_Iwinv"] == rirt(Ibinv)
    Iw = np.array(got["box_Iw"]).reshape(3, 3)
    assert np.allclose(Iw, Iw.T, rtol=0, atol=1e-12) and abs(np.trace(Iw) - np.trace(Ib)) < 1e-12 * np.trace(Ib)
    assert got["world"] == [0.0, 0.0, -9.81, 0.875, 2.0]   # getBody(1) is the sphere of uid 2
    bp = (0.25, -0.5, 0.125)
    assert got["box_point"] == [pos[k] + (R[k, 0] * bp[0] + R[k, 1] * bp[1] + R[k, 2] * bp[2]) for k in range(3)]
    # sphere, capsule (Euler angles turn the body's own quaternion)
    check("sph1", np.array([-1, 0.75, 1.5]), mul(Q((1, 0, 1), 0.6), np.array([1.0, 0, 0, 0])))
    qc = mul(Q((0, 0, 1), 0.25), mul(Q((0, 1, 0), -0.1), mul(Q((1, 0, 0), 0.7), np.array([1.0, 0, 0, 0]))))
    check("cap1", np.array([0.5, 2, 3]), qc)
    # plane: the normal is R's third column; its displacement n.gpos, kept by rotateAroundOrigin
    prow = np.array(got["pl0"])
    pos, q, d = prow[:3], prow[3:7], -0.5

    def nrm(q):
        R = O.quat_to_rotation(q)
        return np.array([R[0, 2], R[1, 2], R[2, 2]])

    def disp(n, p):
        return n[0] * p[0] + n[1] * p[1] + n[2] * p[2]
    check("pl0", pos, q, (0, 0, 1), d)
    q = mul(Q((1, 0.5, 0), 0.3), q); n = nrm(q); d = disp(n, pos); check("pl1", pos, q, n, d)
    dq = Q((0, 1, 0.2), 0.25); pos = O.quat_rotate(dq, pos); q = mul(dq, q); n = nrm(q); check("pl2", pos, q, n, d)
    pos = np.array([0.2, 0.3, 1.5]); d = disp(n, pos); check("pl3", pos, q, n, d)
    q = Q((1, 1, 0), 0.3); n = nrm(q); d = disp(n, pos); check("pl4", pos, q, n, d)
    pos = np.array([pos[0] + 0.05, pos[1] + -0.1, pos[2] + 0.2]); d = disp(n, pos); check("pl5", pos, q, n, d)
    # union: its members follow (gpos + R rpos after a turn, + dp after a translation)
    pos, q = np.array([2.0, 2.0, 3.0]), np.array([1.0, 0, 0, 0])
    mem = np.array(got["un0"][11:]).reshape(-1, 3)
    check("un0", pos, q, mem=mem)
    q = mul(euler(0.3, 0.6, -0.4), q); mem = members(pos, q); check("un1", pos, q, mem=mem)
    dq = Q((0.5, -1, 0.3), 0.45); q = mul(dq, q); pos = O.quat_rotate(dq, pos); mem = members(pos, q)
    check("un2", pos, q, mem=mem)
    dp = (-0.2, 0.1, 0.35)
    pos = np.array([pos[k] + dp[k] for k in range(3)]); mem = mem + np.array(dp); check("un3", pos, q, mem=mem)
    dq = Q((0, 0, 1), -0.8); q = mul(dq, q); pos = turned_about(pos, dq, (1, 1, 1)); mem = members(pos, q)
    check("un4", pos, q, mem=mem)

    # pe's random numbers (Random.h): default seed 12345; setSeed repeats the sequence, and the
    # granular particle drawn after it; rand<real>() in [0, 1), rand<real>(-2, 3) in [-2, 3),
    # rand<int>(1, 6) in [1, 6]
    assert got["seed0"] == [12345]
    a, b = got["draws0"], got["draws1"]
    assert a == b and a[0] == 7
    assert all(0 <= x < 1 for x in a[1:5]) and all(-2 <= x < 3 for x in a[5:9])
    assert all(x in (1, 2, 3, 4, 5, 6) for x in a[9:13])
    n = int(a[13])
    assert 2 <= n <= 4 and len(a) == 14 + n and all(0.3 <= r <= 0.4 for r in a[14:])
    # timing::WcTimer: count, total = the sum, average, min <= last <= max, all positive
    cnt, total, avg, tmin, tmax, last = got["timer"]
    assert cnt == 3 and 0 < tmin <= last <= tmax and tmin <= avg <= tmax
    assert abs(total - 3 * avg) <= 1e-12 * total


def test_cpp_dropin_host_code_clean_under_sanitizers(tmp_path):
    # The drop-in's host code (body factories, union setters, destroy, iterators, pose utilities,
    # random numbers, body files) built with AddressSanitizer and UndefinedBehaviorSanitizer by g++ (no
    # recovery, leak detection on): both test programs run without a report and print what the
    # plain builds print (the wall-clock timer line aside).
    import shutil
    import subprocess
    if shutil.which("g++") is None or not os.path.exists(_abi.LIB_PATH):
        pytest.skip("g++ or the library not available")
    lib = os.path.dirname(_abi.LIB_PATH)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:exitcode=99", UBSAN_OPTIONS="print_stacktrace=1")
    for name in ("dropin_host_arrays", "dropin_rotations", "dropin_body_file"):
        src = os.path.join(ROOT, "tests", "native", name + ".cpp")
        outs = []
        for flags in ([], ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]):
            exe = str(tmp_path / (name + ("_san" if flags else "")))
            subprocess.run(["g++", "-O2", "-std=c++17", *flags, "-I", os.path.join(ROOT, "include"), src, "-o", exe,
                            "-L", lib, "-lpehip", "-Wl,-rpath," + lib], check=True)
            runs = [[exe]]
            if name == "dropin_body_file":   # write a file, then read it back
                f = str(tmp_path / ("bf_san.pe" if flags else "bf.pe"))
                runs = [[exe, "write", f, "0"], [exe, "read", f]]
            lines = []
            for cmd in runs:
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, env=env)
                assert r.returncode == 0 and r.stderr == "", (name, flags, r.stderr[-2000:])
                lines += [l for l in r.stdout.splitlines() if not l.startswith("timer")]
            outs.append(lines)
        assert outs[0] == outs[1] and outs[0], name


def _mixed_world(World, quat_axis_angle):
    # the scene of tests/native/dropin_body_file.cpp (write mode), built with the Python World
    w = World()
    gran = w.createMaterial("granular", 2500, 0, 0.2, 0.2)
    wood = w.createMaterial("wood", 700, 0.1, 0.35, 0.3)
    w.createPlane(100, (0, 0, 1), 0, gran)
    w.createPlane(101, (1, -2, 0.5), -3.25, wood)
    s1 = w.createSphere(1, (1.5, -0.25, 2), 0.375, gran)
    box = w.createBox(2, (-1, 2, 0.75), (0.5, 1.25, 0.8), wood)
    w.createCapsule(3, (0.3, 0.1, 4), 0.2, 1.1, gran)
    u = w.createUnion(4, (2, 2, 3), [((0.3, 0, 0), 0.25, gran), ((-0.1, 0.2, 0.05), 0.3, wood),
                                     ((0, -0.15, 0.2), 0.2, gran)])
    w.createUnion(5, (-2, -1, 1.5), [((0.1, 0.1, 0.1), 0.4, wood)])
    w.createSphere(6, (0, 0, 7), 0.5, wood)
    w.setFixed(box)
    a = w.host_arrays()
    a["v"][s1] = (0.5, -0.25, 1.0)
    a["w"][s1] = (2, 0, -1)
    a["q"][s1] = quat_axis_angle((1, 2, 3), 0.4)
    a["v"][u] = (0, 0, -2)
    return w


def test_cpp_dropin_body_files_equal_python_ones(tmp_path):
    # BodyBinaryWriter / BodyBinaryReader in the C++ drop-in against pe_amd.checkpoint (pe's format,
    # pinned by tests/test_checkpoint.py): the same scene written by both is the same bytes, in
    # double and in single precision, and a file read by both gives the same body and member
    # arrays, system IDs and next system ID (pe's reader order: local chunk, then global).
    import shutil
    import subprocess
    from pe_amd import checkpoint
    from pe_amd.world import quat_axis_angle
    if shutil.which("g++") is None or not os.path.exists(_abi.LIB_PATH):
        pytest.skip("g++ or the library not available")
    exe = str(tmp_path / "dropin_body_file")
    lib = os.path.dirname(_abi.LIB_PATH)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "dropin_body_file.cpp"), "-o", exe, "-L", lib, "-lpehip",
                    "-Wl,-rpath," + lib], check=True)
    for fp in (0, 2):
        fc, fpy = str(tmp_path / ("c%d.pe" % fp)), str(tmp_path / ("py%d.pe" % fp))
        subprocess.run([exe, "write", fc, str(fp)], check=True, timeout=60)
        checkpoint.write_body_file(_mixed_world(World, quat_axis_angle), fpy, fp_size=fp)
        assert open(fc, "rb").read() == open(fpy, "rb").read(), fp

        out = subprocess.run([exe, "read", fpy], capture_output=True, text=True, check=True, timeout=60).stdout
        cpp = {l.split()[0]: np.array([float(x) for x in l.split()[1:]]) for l in out.splitlines()}
        w = World()
        w.createMaterial("granular", 2500, 0, 0.2, 0.2)
        w.createMaterial("wood", 700, 0.1, 0.35, 0.3)
        checkpoint.load_body_file(w, fpy)
        a, m = w.host_arrays(), w.members()
        for k in ("shape", "sid", "uid", "mat", "fixed", "param", "pos", "q", "v", "w", "inv_mass", "I", "Iinv",
                  "normal", "d"):
            assert np.array_equal(cpp[k], np.asarray(a[k], float).ravel()), (fp, k)
        for k, mk in (("m_owner", "owner"), ("m_sid", "sid"), ("m_uid", "uid"), ("m_mat", "mat"),
                      ("m_rad", "radius"), ("m_rpos", "rpos"), ("m_gpos", "gpos")):
            assert np.array_equal(cpp[k], np.asarray(m[mk], float).ravel()), (fp, k)
        assert int(cpp["next_sid"][0]) == w._next_sid
        assert cpp["sid"].tolist() == [3, 12, 4, 5, 6, 10, 1, 2]   # local chunk by type, then the planes
    # a file that is refused leaves the world empty: bodies of a material the world lacks, or the
    # file cut short (the whole file is checked before the first body is created)
    import struct
    fpy = str(tmp_path / "py0.pe")
    cut, short = str(tmp_path / "cut.pe"), str(tmp_path / "short.pe")
    raw = open(fpy, "rb").read()[:-20]
    open(cut, "wb").write(raw)                       # offsets past the end of the file
    open(short, "wb").write(raw[:21] + struct.pack(">I", len(raw)) + raw[25:])   # a record cut short
    for f, nmat, why in ((fpy, "1", "which the world does not have"), (cut, "2", "Invalid body data offset"),
                         (short, "2", "truncated"), (fpy, "2", None)):
        r = subprocess.run([exe, "tryread", f, nmat], capture_output=True, text=True, check=True, timeout=60).stdout
        if why:
            assert r.startswith("refused") and why in r and "bodies 0 members 0" in r, r
        else:
            assert r.startswith("read") and "bodies 8 members 4" in r, r
