"""pe binary body files (BodyBinaryWriter / BodyBinaryReader format): round trips through
pe_amd.checkpoint, the byte layout fixed by the reference's writer, and restarts on the GPU."""
import math
import struct

import numpy as np
import pytest

import oracle as O
from pe_amd import World, checkpoint, scenes
from pe_amd._abi import BOX, CAPSULE, PLANE, SPHERE, UNION


def mixed_world():
    rng = np.random.default_rng(3)
    w = World()
    w.setGravity(0, 0, -9.81)
    w.createMaterial("iron", 7800.0, 0.2, 0.3, 0.25)
    w.createMaterial("wood", 600.0, 0.4, 0.2, 0.1)
    w.createPlane(11, (0, 0, 1), 0.0, 0)
    w.createPlane(12, (0, 0, -1), -20.0, 0)
    w.createPlane(13, (0.5, 0.1, math.cos(math.radians(30))), 0.3, 1)
    for k in range(12):
        w.createSphere(100 + k, rng.uniform(1, 9, 3), rng.uniform(0.2, 0.6), k % 2)
    for k in range(5):
        w.createBox(200 + k, rng.uniform(1, 9, 3), rng.uniform(0.3, 0.9, 3), 1 - k % 2)
    w.setFixed(5)
    a = w.host_arrays()
    n = len(a["shape"])
    fin = a["shape"] != PLANE
    a["v"][fin] = rng.uniform(-1, 1, (fin.sum(), 3))
    a["w"][fin] = rng.uniform(-2, 2, (fin.sum(), 3))
    boxes = a["shape"] == BOX
    q = scenes.shoemake_quaternions(rng, int(boxes.sum()))
    a["q"][boxes] = q
    a["v"][5] = 0.0
    a["w"][5] = 0.0
    assert n == 20
    return w


def pe_reader_order(a):
    """Body indices in the order pe's reader creates them: spheres, boxes, capsules (local), planes
    (global)."""
    s = a["shape"]
    return np.concatenate([np.flatnonzero(s == SPHERE), np.flatnonzero(s == BOX), np.flatnonzero(s == CAPSULE),
                           np.flatnonzero(s == PLANE)])


def test_round_trip_reproduces_every_body(tmp_path):
    w = mixed_world()
    f = tmp_path / "bodies.bin"
    checkpoint.write_body_file(w, str(f))
    w2 = World()
    for m in w.materials:
        w2.createMaterial(m.name, m.density, m.cor, m.csf, m.cdf)
    info = checkpoint.load_body_file(w2, str(f))
    a, b = w.host_arrays(), w2.host_arrays()
    order = pe_reader_order(a)
    assert len(info["bodies"]) == len(order) == len(b["shape"])
    for k in ("shape", "sid", "uid", "mat", "fixed", "pos", "q", "v", "w", "param", "inv_mass", "I", "Iinv"):
        assert np.array_equal(np.asarray(a[k])[order], np.asarray(b[k])), k
    assert w2._next_sid == w._next_sid
    # a read plane takes its normal from its orientation (Plane.cpp:121-142), as pe's reader
    # does, not the exact normal the plane was created with: the -z plane gains a -1.2e-16
    for i in np.flatnonzero(b["shape"] == PLANE):
        nrm = checkpoint.quat_rotate(b["q"][i], (0.0, 0.0, 1.0))
        assert np.array_equal(b["normal"][i], nrm)
        assert b["d"][i] == nrm[0] * b["pos"][i][0] + nrm[1] * b["pos"][i][1] + nrm[2] * b["pos"][i][2]
        assert np.abs(b["normal"][i] - np.asarray(a["normal"])[order][i]).max() < 1e-15


def test_plane_orientation_gives_pe_normal(tmp_path):
    # the reader's plane normal is q.rotate(0, 0, 1) (Plane.cpp:121-142): the oracle's
    # restatement of Quaternion::rotate gives the same bits, including the -z plane's artefact
    w = mixed_world()
    a = w.host_arrays()
    for i in np.flatnonzero(a["shape"] == PLANE):
        nrm = checkpoint.quat_rotate(a["q"][i], (0.0, 0.0, 1.0))
        assert np.array_equal(np.array(nrm), O.quat_rotate(a["q"][i], np.array([0.0, 0.0, 1.0])))
    nz = checkpoint.quat_rotate(a["q"][1], (0.0, 0.0, 1.0))
    assert nz[2] == -1.0 and nz[1] != 0.0           # acos(-1) = pi: cos(pi/2) is not 0 in binary


def test_byte_layout(tmp_path):
    w = World()
    w.createMaterial("m", 1000.0, 0.0, 0.1, 0.1)
    w.createPlane(7, (0, 0, 1), 0.0, 0)
    w.createSphere(9, (1.0, 2.0, 3.0), 0.5, 0)
    f = tmp_path / "one.bin"
    checkpoint.write_body_file(w, str(f))
    data = f.read_bytes()
    # header: magic, version 0.1, sizes of real/int/size_t/id_t/bool, one process, offsets
    assert data[:9] == bytes([ord("P"), ord("E"), 0, 1, 8, 4, 8, 8, 1])
    p, off_g, off_l, end = struct.unpack(">IIII", data[9:25])
    assert p == 1 and off_g == 25 and end == len(data)
    # global chunk: counter, then (type byte, uint32 count) for sphere, box, capsule, cylinder,
    # plane (one plane), mesh, union
    g = data[off_g:off_l]
    assert struct.unpack(">Q", g[:8])[0] == 1
    assert g[8:13] == bytes([1, 0, 0, 0, 0]) and g[13:18] == bytes([2, 0, 0, 0, 0])
    assert g[28:33] == bytes([5, 0, 0, 0, 1])
    plane = g[33:33 + 8 + 8 + 2 + 13 * 8 + 8]
    sid, uid, vis, fixed = struct.unpack(">QQBB", plane[:18])
    assert (uid, vis, fixed) == (7, 1, 1)
    # local chunk: counter, then the sphere record; 1.0 as a big-endian double is 3FF0000000000000
    l = data[off_l:]
    assert struct.unpack(">Q", l[:8])[0] == w._next_sid
    assert l[8:13] == bytes([1, 0, 0, 0, 1])
    rec = l[13:]
    assert struct.unpack(">QQBB", rec[:18])[1:] == (9, 1, 0)
    assert rec[18:26] == bytes.fromhex("3ff0000000000000")
    assert struct.unpack(">3d", rec[18:42]) == (1.0, 2.0, 3.0)
    assert struct.unpack(">4d", rec[42:74]) == (1.0, 0.0, 0.0, 0.0)
    radius_at = 18 + 13 * 8 + 8
    assert rec[radius_at:radius_at + 8] == bytes.fromhex("3fe0000000000000")   # 0.5
    assert len(rec) == radius_at + 8 + 5 * 5                  # then box, capsule, cylinder, mesh, union: empty


def test_single_precision_file(tmp_path):
    w = mixed_world()
    f = tmp_path / "single.bin"
    checkpoint.write_body_file(w, str(f), fp_size=2)
    info = checkpoint.read_body_file(str(f))
    a = w.host_arrays()
    order = pe_reader_order(a)
    pos = np.array([b["pos"] for b in info["bodies"]])
    assert np.array_equal(pos, a["pos"][order].astype(np.float32).astype(np.float64))


def test_rejects_bad_files(tmp_path):
    w = mixed_world()
    f = tmp_path / "b.bin"
    checkpoint.write_body_file(w, str(f))
    data = bytearray(f.read_bytes())
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"XE" + bytes(data[2:]))
    with pytest.raises(ValueError, match="Invalid file format"):
        checkpoint.read_body_file(str(bad))
    v = bytearray(data)
    v[3] = 2
    bad.write_bytes(bytes(v))
    with pytest.raises(ValueError, match="Unsupported file format version"):
        checkpoint.read_body_file(str(bad))
    bad.write_bytes(bytes(data[:-7]))
    with pytest.raises(ValueError):
        checkpoint.read_body_file(str(bad))
    # a process count far beyond the file: refused before its offset table is read
    v = bytearray(data)
    v[9:13] = struct.pack(">I", 0xFFFFFFFF)
    bad.write_bytes(bytes(v))
    with pytest.raises(ValueError, match="truncated body file"):
        checkpoint.read_body_file(str(bad))
    # a cylinder (type 4) in the local chunk: not a shape of this library
    off_l = struct.unpack(">I", bytes(data[17:21]))[0]
    v = bytearray(data)
    n_sph = struct.unpack(">I", bytes(v[off_l + 9:off_l + 13]))[0]
    rec = 18 + 13 * 8 + 8 + 8
    box_hdr = off_l + 13 + n_sph * rec
    n_box = struct.unpack(">I", bytes(v[box_hdr + 1:box_hdr + 5]))[0]
    cap_hdr = box_hdr + 5 + n_box * (rec + 16)
    assert v[cap_hdr] == 3 and struct.unpack(">I", bytes(v[cap_hdr + 1:cap_hdr + 5]))[0] == 0
    cyl_hdr = cap_hdr + 5
    assert v[cyl_hdr] == 4
    v[cyl_hdr + 1:cyl_hdr + 5] = struct.pack(">I", 1)
    bad.write_bytes(bytes(v))
    with pytest.raises(ValueError, match="Unsupported geometry type 4"):
        checkpoint.read_body_file(str(bad))


def capsule_world():
    rng = np.random.default_rng(8)
    w = World()
    w.createMaterial("rubber", 1100.0, 0.5, 0.6, 0.5)
    w.createPlane(1, (0, 0, 1), 0.0, 0)
    for k in range(3):
        w.createSphere(10 + k, rng.uniform(1, 5, 3), 0.3, 0)
    for k in range(4):
        w.createCapsule(20 + k, rng.uniform(1, 5, 3), rng.uniform(0.1, 0.4), rng.uniform(0.3, 1.5), 0)
    a = w.host_arrays()
    caps = a["shape"] == CAPSULE
    a["q"][caps] = scenes.shoemake_quaternions(rng, int(caps.sum()))
    a["v"][caps] = rng.uniform(-1, 1, (int(caps.sum()), 3))
    return w


def test_capsule_records_round_trip(tmp_path):
    # capsules are local bodies after spheres and boxes, their record ending in radius and length
    # (Marshalling.h:352-356); reading them back gives pe's capsule mass properties
    w = capsule_world()
    f = tmp_path / "caps.bin"
    checkpoint.write_body_file(w, str(f))
    data = f.read_bytes()
    off_l = struct.unpack(">I", data[17:21])[0]
    n_sph = struct.unpack(">I", data[off_l + 9:off_l + 13])[0]
    rec = 18 + 13 * 8 + 8 + 8
    box_hdr = off_l + 13 + n_sph * rec
    assert data[box_hdr] == 2 and struct.unpack(">I", data[box_hdr + 1:box_hdr + 5])[0] == 0
    cap_hdr = box_hdr + 5
    assert data[cap_hdr] == 3 and struct.unpack(">I", data[cap_hdr + 1:cap_hdr + 5])[0] == 4
    a = w.host_arrays()
    ci = np.flatnonzero(a["shape"] == CAPSULE)
    first = cap_hdr + 5
    crec = rec + 8                                        # radius and length: one real more
    for k, i in enumerate(ci):
        tail = data[first + k * crec + crec - 16:first + k * crec + crec]
        assert struct.unpack(">dd", tail) == (a["param"][i][0], a["param"][i][1])
    w2 = World()
    for m in w.materials:
        w2.createMaterial(m.name, m.density, m.cor, m.csf, m.cdf)
    checkpoint.load_body_file(w2, str(f))
    b = w2.host_arrays()
    order = pe_reader_order(a)
    for key in ("shape", "sid", "uid", "pos", "q", "v", "w", "param", "inv_mass", "I", "Iinv"):
        assert np.array_equal(np.asarray(a[key])[order], np.asarray(b[key])), key


@pytest.mark.gpu
def test_restart_from_file_on_gpu(tmp_path):
    # a run interrupted by a body file continues with the same bodies: same state and the same
    # contact set (up to pe's reader body order) as the world it was written from
    # a nearly touching lattice just above the floor: dozens of contacts by step 60
    sc = scenes.drop_scene(4, 4, 3, r=0.45, spacing=0.902, jitter=0.001, z0=0.451, box=(8.0, 8.0, 20.0),
                           v_amp=0.1, seed=4)
    w = World()
    sc.to_world(w)
    w.run(60, sc.dt)
    f = tmp_path / "mid.bin"
    checkpoint.write_body_file(w, str(f))
    w2 = World()
    for m in w.materials:
        w2.createMaterial(m.name, m.density, m.cor, m.csf, m.cdf)
    w2.setGravity(*sc.gravity)
    checkpoint.load_body_file(w2, str(f))
    a = w.host_arrays()
    order = pe_reader_order(a)
    s1, s2 = w.getState(), w2.getState()
    for k in ("pos", "q", "v", "w", "normal", "d"):
        assert np.array_equal(s1[k][order], s2[k]), k
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    c1, c2 = w.detect(), w2.detect()
    k1 = sorted(zip(inv[c1["b1"]].tolist(), inv[c1["b2"]].tolist(), map(tuple, c1["pos"].tolist())))
    k2 = sorted(zip(c2["b1"].tolist(), c2["b2"].tolist(), map(tuple, c2["pos"].tolist())))
    assert len(k1) > 10 and k1 == k2


def union_world():
    """Spheres, a box, unions (one with its own member user IDs) and planes, with velocities."""
    rng = np.random.default_rng(21)
    w = World()
    g = w.createMaterial("glass", 2500.0, 0.0, 0.2, 0.2)
    s = w.createMaterial("steel", 7800.0, 0.0, 0.3, 0.3)
    w.createPlane(1, (0, 0, 1), 0.0, g)
    w.createSphere(2, (0.0, 0.0, 1.0), 0.5, g)
    w.createBox(3, (3.0, 0.0, 1.0), (0.5, 0.6, 0.7), s)
    w.createUnion(4, (2.0, 0.0, 1.0), [((0.0, 0.0, 0.0), 0.3, g, 40), ((0.4, 0.1, 0.0), 0.2, s, 41)])
    for k in range(3):
        w.createGranularParticle(10 + k, rng.uniform(1, 6, 3), 0.5, g, rng=rng)
    w.createPlane(5, (1, 0, 0), -1.0, s)
    w.createSphere(6, (5.0, 5.0, 1.0), 0.4, s)
    a = w.host_arrays()
    fin = a["shape"] != PLANE
    a["v"][fin] = rng.uniform(-1, 1, (int(fin.sum()), 3))
    a["w"][fin] = rng.uniform(-2, 2, (int(fin.sum()), 3))
    return w


def reader_order(a):
    """pe's reader order with unions: spheres, boxes, capsules, unions (local), planes (global)."""
    s = a["shape"]
    return np.concatenate([np.flatnonzero(s == t) for t in (SPHERE, BOX, CAPSULE, UNION, PLANE)])


def copy_materials(w):
    w2 = World()
    for m in w.materials:
        w2.createMaterial(m.name, m.density, m.cor, m.csf, m.cdf)
    return w2


def test_union_round_trip(tmp_path):
    # unions are local bodies after capsules; reading them back keeps the file's mass, body
    # inertia, orientation, velocities and the members' IDs, materials, radii and positions
    # (BodyBinaryReader.cpp:432-500, 594-603), and writing the read world gives the same bytes
    w = union_world()
    f = tmp_path / "u.bin"
    checkpoint.write_body_file(w, str(f))
    w2 = copy_materials(w)
    checkpoint.load_body_file(w2, str(f))
    a, b = w.host_arrays(), w2.host_arrays()
    order = reader_order(a)
    assert len(order) == len(b["shape"]) == 9
    for k in ("shape", "sid", "uid", "mat", "fixed", "pos", "q", "v", "w", "param", "inv_mass", "I", "Iinv"):
        assert np.array_equal(np.asarray(a[k])[order], np.asarray(b[k])), k
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    m1, m2 = w.members(), w2.members()
    assert np.array_equal(inv[m1["owner"]], m2["owner"])   # unions stay in creation order
    for k in ("sid", "uid", "mat", "radius", "rpos", "gpos"):
        assert np.array_equal(m1[k], m2[k]), k
    assert list(m1["uid"][:2]) == [40, 41] and set(m1["uid"][2:]) == {10, 11, 12}
    for u in np.flatnonzero(a["shape"] == UNION):
        assert w2.unionMass(int(inv[u])) == w.unionMass(int(u))
    assert w2._next_sid == w._next_sid
    f2 = tmp_path / "u2.bin"
    checkpoint.write_body_file(w2, str(f2))
    assert f2.read_bytes() == f.read_bytes()


def test_union_byte_layout(tmp_path):
    # one union of two spheres: its record (Marshalling.h:523-570) ends the local chunk
    w = World()
    m = w.createMaterial("m", 1000.0, 0.0, 0.1, 0.1)
    w.createUnion(9, (1.0, 2.0, 3.0), [((0.0, 0.0, 0.0), 0.5, m), ((1.0, 0.0, 0.0), 0.25, m, 77)])
    a = w.host_arrays()
    a["v"][0] = (0.5, -0.25, 2.0)
    f = tmp_path / "one_union.bin"
    checkpoint.write_body_file(w, str(f))
    data = f.read_bytes()
    off_l = struct.unpack(">I", data[17:21])[0]
    l = data[off_l:]
    assert struct.unpack(">Q", l[:8])[0] == 4                 # the union and its two spheres: IDs 1-3
    # sphere, box, capsule, cylinder and mesh blocks are empty; then one union
    assert l[8:33] == bytes([1, 0, 0, 0, 0, 2, 0, 0, 0, 0, 3, 0, 0, 0, 0, 4, 0, 0, 0, 0, 6, 0, 0, 0, 0])
    assert l[33:38] == bytes([7, 0, 0, 0, 1])
    r = l[38:]
    assert struct.unpack(">QQBB", r[:18]) == (1, 9, 1, 0)
    x = struct.unpack(">13d", r[18:18 + 104])
    assert x[:3] == (1.0, 2.0, 3.0) and x[3:7] == (1.0, 0.0, 0.0, 0.0) and x[7:10] == (0.5, -0.25, 2.0)
    o = 18 + 104
    # closed forms: masses 4/3 pi r^3 rho; the centre of mass a ninth of the way to the small
    # sphere (mass ratio 8:1); parallel axes about it
    m1, m2 = 4.0 / 3.0 * math.pi * 0.125 * 1000.0, 4.0 / 3.0 * math.pi * 0.015625 * 1000.0
    mass, = struct.unpack(">d", r[o:o + 8])
    assert mass == w.unionMass(0) and abs(mass - (m1 + m2)) < 1e-12 * mass
    I = struct.unpack(">9d", r[o + 8:o + 80])
    d1, d2 = 1.0 / 9.0, 8.0 / 9.0
    ixx = 0.4 * m1 * 0.25 + 0.4 * m2 * 0.0625
    iyy = ixx + m1 * d1 * d1 + m2 * d2 * d2
    assert abs(I[0] - ixx) < 1e-12 * ixx and abs(I[4] - iyy) < 1e-12 * iyy and abs(I[8] - iyy) < 1e-12 * iyy
    assert I[1] == I[2] == I[3] == I[5] == I[6] == I[7] == 0.0
    aabb = struct.unpack(">6d", r[o + 80:o + 128])
    g1 = 1.0 - d1                                           # the big sphere's x after placement
    exp = (g1 - 0.5 - 1e-8, 2.0 - 0.5 - 1e-8, 3.0 - 0.5 - 1e-8, g1 + 1.0 + 0.25 + 1e-8, 2.0 + 0.5 + 1e-8,
           3.0 + 0.5 + 1e-8)
    assert np.abs(np.array(aabb) - exp).max() < 1e-14
    assert struct.unpack(">Q", r[o + 128:o + 136])[0] == 2
    s = r[o + 136:]
    rec = 1 + 18 + 10 * 8 + 8 + 8                           # type, ids and flags, 3+3+4 reals, material, radius
    assert len(s) == 2 * rec
    mem = w.members()
    for k in range(2):
        t = s[k * rec:(k + 1) * rec]
        assert t[0] == 1
        assert struct.unpack(">QQBB", t[1:19]) == (2 + k, (9, 77)[k], 1, 0)
        v = struct.unpack(">10d", t[19:99])
        assert v[:3] == tuple(mem["gpos"][k]) and v[3:6] == tuple(mem["rpos"][k]) and v[6:] == (1.0, 0.0, 0.0, 0.0)
        assert struct.unpack(">Qd", t[99:115]) == (m, (0.5, 0.25)[k])
    assert abs(mem["rpos"][0][0] + d1) < 1e-15 and abs(mem["rpos"][1][0] - d2) < 1e-15


def test_union_records_rejected(tmp_path):
    w = World()
    m = w.createMaterial("m", 1000.0, 0.0, 0.1, 0.1)
    w.createUnion(9, (1.0, 2.0, 3.0), [((0.0, 0.0, 0.0), 0.5, m), ((1.0, 0.0, 0.0), 0.25, m)])
    f = tmp_path / "u.bin"
    checkpoint.write_body_file(w, str(f))
    data = bytearray(f.read_bytes())
    off_l = struct.unpack(">I", bytes(data[17:21]))[0]
    urec = off_l + 38
    first_member = urec + 18 + 104 + 136
    bad = tmp_path / "bad.bin"
    v = bytearray(data)
    v[first_member] = 2                                      # a box member
    bad.write_bytes(bytes(v))
    with pytest.raises(ValueError, match="Unsupported geometry type 2 in a union"):
        checkpoint.read_body_file(str(bad))
    v = bytearray(data)
    v[urec + 17] = 1                                         # a fixed union
    bad.write_bytes(bytes(v))
    with pytest.raises(ValueError, match="fixed unions"):
        checkpoint.load_body_file(copy_materials(w), str(bad))
    w3 = World()                                             # a member material the world lacks
    with pytest.raises(ValueError, match="material 0"):
        checkpoint.load_body_file(w3, str(f))


def test_union_single_precision_file(tmp_path):
    # reals of 4 bytes (setFloatingPointSize(2)) in union records too: the mass, inertia, bounding
    # box and the members' positions read back as the float32 roundings of the world's values
    w = union_world()
    f = tmp_path / "u32.bin"
    checkpoint.write_body_file(w, str(f), fp_size=2)
    info = checkpoint.read_body_file(str(f))
    unions = [b for b in info["bodies"] if b["type"] == checkpoint.UNION_TYPE]
    a = w.host_arrays()
    ui = np.flatnonzero(a["shape"] == UNION)
    m = w.members()
    assert len(unions) == len(ui) == 4
    for b, i in zip(unions, ui):
        f32 = lambda x: np.asarray(x, np.float64).astype(np.float32).astype(np.float64)
        assert np.array_equal(b["pos"], f32(a["pos"][i])) and np.array_equal(b["I"], f32(a["I"][i]))
        assert b["mass"] == f32(w.unionMass(int(i)))
        sel = np.flatnonzero(m["owner"] == i)
        assert [mm["sid"] for mm in b["members"]] == [int(x) for x in m["sid"][sel]]
        assert np.array_equal(np.array([mm["pos"] for mm in b["members"]]), f32(m["gpos"][sel]))


def test_bulk_unions_are_checked():
    # add_bodies with union rows needs their members, grouped by union in body order
    w = union_world()
    a = {k: np.array(v, copy=True) for k, v in w.host_arrays().items()}
    m = w.members()
    for bad, msg in ((None, "members"),
                     (dict(m, owner=m["owner"][::-1].copy()), "grouped"),
                     ({k: v[m["owner"] != m["owner"][0]] for k, v in m.items()}, "without members")):
        w2 = copy_materials(w)
        with pytest.raises(ValueError, match=msg):
            w2.add_bodies(a, sids=a["sid"], members=bad)


@pytest.mark.gpu
def test_union_restart_continues_bit_for_bit(tmp_path):
    # granular particles above a floor, created in pe's reader order (unions, then planes): a
    # world read back from a body file written mid-run continues exactly as the original
    rng = np.random.default_rng(5)
    w = World()
    g = w.createMaterial("granular", 2500.0, 0.0, 0.3, 0.3)
    w.setGravity(0.0, 0.0, -9.81)
    for k in range(60):
        i, j, l = k % 5, (k // 5) % 4, k // 20
        w.createGranularParticle(k, (0.6 + 1.02 * i, 0.6 + 1.02 * j, 0.55 + 1.02 * l), 0.5, g, rng=rng)
    for nrm, d in (((0, 0, 1), 0.0), ((1, 0, 0), 0.0), ((-1, 0, 0), -5.2), ((0, 1, 0), 0.0), ((0, -1, 0), -4.2)):
        w.createPlane(100, nrm, d, g)
    w.run(200, 2e-3)                   # landed on the floor and on each other
    f = tmp_path / "mid.bin"
    checkpoint.write_body_file(w, str(f))
    w2 = copy_materials(w)
    w2.setGravity(0.0, 0.0, -9.81)
    checkpoint.load_body_file(w2, str(f))
    a = w.host_arrays()
    assert np.array_equal(reader_order(a), np.arange(len(a["shape"])))
    assert w.getNumberOfContacts() > 15
    w.run(40, 2e-3)
    w2.run(40, 2e-3)
    s1, s2 = w.getState(), w2.getState()
    for k in ("pos", "q", "v", "w", "mgpos"):
        assert np.array_equal(s1[k], s2[k]), k
