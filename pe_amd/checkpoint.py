"""pe's binary rigid-body files: the format of ``BodyBinaryWriter`` / ``BodyBinaryReader``.

Layout (src/core/BodyBinaryWriter.cpp:46-175, src/core/BodyBinaryReader.cpp:60-330,
pe/core/Marshalling.h:55-500), every value big-endian (``HtonEndiannessConversion``,
pe/util/Endianness.h:308):

    header   'P' 'E' 0 1                       magic, format version 0.1
             sizeof(real) int size_t id_t bool  one byte each: 8 4 8 8 1
             uint32 p                           writing processes
             uint32 offset of the global chunk
             uint32 offset of the first local chunk
             uint32 end of local chunk i, i < p
    globals  id_t global ID counter, then per geometry type (sphere, box, capsule, cylinder,
             plane, triangle mesh, union): byte type, uint32 count, the bodies
    locals   id_t ID counter, then the same for sphere, box, capsule, cylinder, triangle mesh,
             union (planes only ever in the global chunk, BodyBinaryWriter.cpp:57)

A top-level body is sid (id_t), uid (id_t), visible (byte), fixed (byte), position (3 reals),
orientation quaternion (4), linear and angular velocity (3 + 3), material (size_t), then the
shape's parameters: a sphere's radius, a box's three edge lengths, a capsule's radius and
length, nothing for a plane (Marshalling.h:184-210, 290-300, 321-330, 352-356, 416-420, 498-501).  Reals are doubles unless the file
was written with a smaller floating-point size (the header's first size byte: 4 = single).

A union (Marshalling.h:523-570) is the common record without a material, then its total mass
(real), its body-frame inertia (9 reals, row by row), its bounding box (xmin, ymin, zmin, xmax,
ymax, zmax), its number of members (size_t) and per member a type byte and the member's record:
a member carries its position relative to the union after its global one and no velocities
(Marshalling.h:184-197), so a sphere member is sid, uid, visible, fixed, position (3), relative
position (3), orientation (4), material, radius.

The reader creates the local chunk's bodies first, then the global chunk's
(BodyBinaryReader.cpp:265-268), keeping their system IDs; a plane's normal is its orientation
applied to +z and its displacement the normal's dot product with its position
(src/core/rigidbody/Plane.cpp:121-142); a union keeps the file's mass, inertia and orientation and
its members their positions (BodyBinaryReader.cpp:432-500, 594-603; Union.cpp:100-133).
Cylinders, meshes, unions of anything but spheres and fixed unions are not bodies of this
library: reading a file that holds any raises ``ValueError``.
"""
import math
import struct

import numpy as np

from ._abi import BOX, CAPSULE, PLANE, SPHERE, UNION
from .world import CONTACT_THRESHOLD, mat3_inverse

# pe/core/GeomType.h:45-53
SPHERE_TYPE, BOX_TYPE, CAPSULE_TYPE, CYLINDER_TYPE, PLANE_TYPE, MESH_TYPE, UNION_TYPE = 1, 2, 3, 4, 5, 6, 7
_LOCAL_ORDER = (SPHERE_TYPE, BOX_TYPE, CAPSULE_TYPE, CYLINDER_TYPE, MESH_TYPE, UNION_TYPE)
_GLOBAL_ORDER = (SPHERE_TYPE, BOX_TYPE, CAPSULE_TYPE, CYLINDER_TYPE, PLANE_TYPE, MESH_TYPE, UNION_TYPE)
_SHAPE_OF = {SPHERE_TYPE: SPHERE, BOX_TYPE: BOX, CAPSULE_TYPE: CAPSULE, PLANE_TYPE: PLANE}
_TYPE_OF = {SPHERE: SPHERE_TYPE, BOX: BOX_TYPE, CAPSULE: CAPSULE_TYPE, PLANE: PLANE_TYPE}
# reals after the common record: radius (Marshalling.h:296-300); lengths (:318-322); radius,
# length (:352-356); none for a plane
_NPARAM = {SPHERE_TYPE: 1, BOX_TYPE: 3, CAPSULE_TYPE: 2, PLANE_TYPE: 0}


def quat_rotate(q, v):
    """Quaternion::rotate (pe/math/Quaternion.h:744-756), pe's evaluation order."""
    w = q[1] * v[0] + q[2] * v[1] + q[3] * v[2]
    x = q[0] * v[0] - q[3] * v[1] + q[2] * v[2]
    y = q[0] * v[1] - q[1] * v[2] + q[3] * v[0]
    z = q[0] * v[2] - q[2] * v[0] + q[1] * v[1]
    return (q[0] * x + q[1] * w + q[2] * z - q[3] * y,
            q[0] * y + q[2] * w + q[3] * x - q[1] * z,
            q[0] * z + q[3] * w + q[1] * y - q[2] * x)


class _Out:
    def __init__(self, real):
        self.parts = []
        self.size = 0
        self.real = real           # struct code of a real: 'd' or 'f'

    def put(self, fmt, *vals):
        b = struct.pack(">" + fmt, *vals)
        self.parts.append(b)
        self.size += len(b)

    def reals(self, vals):
        self.put(self.real * len(vals), *[float(v) for v in vals])

    def bytes(self):
        return b"".join(self.parts)


class _In:
    def __init__(self, data, real="d"):
        self.data = data
        self.pos = 0
        self.real = real

    def get(self, fmt):
        n = struct.calcsize(">" + fmt)
        if self.pos + n > len(self.data):
            raise ValueError("truncated body file")
        v = struct.unpack_from(">" + fmt, self.data, self.pos)
        self.pos += n
        return v

    def reals(self, n):
        return self.get(self.real * n)


def _marshal_bodies(out, a, idx, gtype, uid):
    out.put("B", gtype)
    out.put("I", len(idx))
    for i in idx:
        out.put("QQBB", int(a["sid"][i]), int(uid[i]), 1, int(a["fixed"][i]))
        out.reals(a["pos"][i])
        out.reals(a["q"][i])
        out.reals(a["v"][i])
        out.reals(a["w"][i])
        out.put("Q", int(a["mat"][i]))
        out.reals(a["param"][i][:_NPARAM[gtype]])


def union_aabb(gpos, radius):
    """A union's bounding box: the merge of its spheres' boxes, each the sphere's centre -/+ its
    radius plus the contact threshold (SphereBase.h:208-218, UnionBase.cpp:209-231)."""
    lo = [math.inf] * 3
    hi = [-math.inf] * 3
    for g, r in zip(gpos, radius):
        length = float(r) + CONTACT_THRESHOLD
        for j in range(3):
            lo[j] = min(lo[j], float(g[j]) - length)
            hi[j] = max(hi[j], float(g[j]) + length)
    return lo + hi


def _marshal_unions(out, world, a, idx, uid):
    """Unions of spheres as Marshalling.h:523-570 writes them: a member's orientation is its
    union's (the spheres are created unrotated inside the union and turn with it)."""
    out.put("B", UNION_TYPE)
    out.put("I", len(idx))
    if not len(idx):
        return
    m = world.members()
    owner = m["owner"]
    for i in idx:
        fixed = int(a["fixed"][i])
        out.put("QQBB", int(a["sid"][i]), int(uid[i]), 1, fixed)
        out.reals(a["pos"][i])
        out.reals(a["q"][i])
        out.reals(a["v"][i])
        out.reals(a["w"][i])
        out.reals([world.unionMass(int(i))])
        out.reals(a["I"][i])
        sel = np.flatnonzero(owner == i)
        out.reals(union_aabb(m["gpos"][sel], m["radius"][sel]))
        out.put("Q", len(sel))
        for k in sel:
            out.put("B", SPHERE_TYPE)
            out.put("QQBB", int(m["sid"][k]), int(m["uid"][k]), 1, fixed)
            out.reals(m["gpos"][k])
            out.reals(m["rpos"][k])
            out.reals(a["q"][i])
            out.put("Q", int(m["mat"][k]))
            out.reals([m["radius"][k]])


def write_body_file(world, filename, fp_size=0):
    """Writes the world's bodies as pe's BodyBinaryWriter::writeFile does for one process
    (planes as global bodies, spheres, boxes, capsules and unions as local ones).  fp_size: 0 =
    doubles (the size of ``real``), 2 = single precision, 3 = double precision
    (setFloatingPointSize)."""
    if fp_size not in (0, 2, 3):
        raise ValueError("Invalid floating point conversion size.")
    real = "f" if fp_size == 2 else "d"
    a = world.host_arrays()
    n = len(a["shape"])
    uid = a.get("uid", np.zeros(n, np.uint64))
    shape = np.asarray(a["shape"])
    planes = np.flatnonzero(shape == PLANE)
    glob = _Out(real)
    glob.put("Q", len(planes))                     # global ID counter
    for t in _GLOBAL_ORDER:
        _marshal_bodies(glob, a, planes if t == PLANE_TYPE else [], t, uid)
    loc = _Out(real)
    loc.put("Q", int(world._next_sid))              # ID counter
    for t in _LOCAL_ORDER:
        if t == UNION_TYPE:
            _marshal_unions(loc, world, a, np.flatnonzero(shape == UNION), uid)
            continue
        idx = np.flatnonzero(shape == _SHAPE_OF[t]) if t in _SHAPE_OF else []
        _marshal_bodies(loc, a, idx, t, uid)
    head = _Out(real)
    head.put("BBBB", ord("P"), ord("E"), 0, 1)
    head.put("BBBBB", 8 if fp_size == 0 else 1 << fp_size, 4, 8, 8, 1)
    head.put("I", 1)
    head_size = head.size + (1 + 2) * 4
    offset = head_size + glob.size
    head.put("III", head_size, offset, offset + loc.size)
    with open(filename, "wb") as f:
        f.write(head.bytes())
        f.write(glob.bytes())
        f.write(loc.bytes())


def read_body_file(filename):
    """Parses a pe binary body file.  Returns {"counter", "global_counter", "bodies"} with the
    bodies in the order pe's reader creates them (local chunks, then the global chunk), each a
    dict: type, sid, uid, visible, fixed, pos, q, v, w, material, param, global; a union has
    mass, I, aabb and members (each: sid, uid, visible, fixed, pos, rpos, q, material, radius)
    in place of material and param."""
    with open(filename, "rb") as f:
        data = f.read()
    h = _In(data)
    m1, m2, major, minor = h.get("BBBB")
    if m1 != ord("P") or m2 != ord("E"):
        raise ValueError("Invalid file format.")
    if (major, minor) != (0, 1):
        raise ValueError("Unsupported file format version.")
    sizes = h.get("BBBBB")
    if sizes[0] not in (4, 8):
        raise ValueError("Invalid size of floating point data.")
    if sizes[1:] != (4, 8, 8, 1):
        raise ValueError("Size of the int, size_t, id_t or bool data type does not match the file.")
    real = "d" if sizes[0] == 8 else "f"
    p, = h.get("I")
    off_global, = h.get("I")
    if h.pos + 4 * (p + 1) > len(data):             # before building a format of p + 1 fields
        raise ValueError("truncated body file")
    offs = h.get("I" * (p + 1))
    if offs[0] < off_global or any(b < a for a, b in zip(offs, offs[1:])) or offs[-1] > len(data):
        raise ValueError("Invalid body data offset.")

    def chunk(lo, hi, is_global):
        r = _In(data[lo:hi], real)
        counter, = r.get("Q")
        out = []
        while r.pos < hi - lo:
            t, = r.get("B")
            cnt, = r.get("I")
            if t == UNION_TYPE:
                out += [dict(_unmarshal_union(r), **{"global": is_global}) for _ in range(cnt)]
                continue
            if cnt and t not in _NPARAM:
                raise ValueError("Unsupported geometry type %d in body file" % t)
            for _ in range(cnt):
                sid, uid, vis, fixed = r.get("QQBB")
                pos, q, v, w = r.reals(3), r.reals(4), r.reals(3), r.reals(3)
                mat, = r.get("Q")
                out.append({"type": t, "sid": sid, "uid": uid, "visible": bool(vis), "fixed": bool(fixed),
                            "pos": pos, "q": q, "v": v, "w": w, "material": mat,
                            "param": r.reals(_NPARAM[t]), "global": is_global})
        return counter, out

    bodies = []
    counter = 0
    for i in range(p):
        c, b = chunk(offs[i], offs[i + 1], False)
        counter = max(counter, c)
        bodies += b
    gcounter, b = chunk(off_global, offs[0], True)
    bodies += b
    return {"counter": counter, "global_counter": gcounter, "bodies": bodies}


def _unmarshal_union(r):
    """One union record (Marshalling.h:560-620), members restricted to spheres."""
    sid, uid, vis, fixed = r.get("QQBB")
    pos, q, v, w = r.reals(3), r.reals(4), r.reals(3), r.reals(3)
    mass, = r.reals(1)
    I, aabb = r.reals(9), r.reals(6)
    size, = r.get("Q")
    members = []
    for _ in range(size):
        mt, = r.get("B")
        if mt != SPHERE_TYPE:
            raise ValueError("Unsupported geometry type %d in a union" % mt)
        msid, muid, mvis, mfixed = r.get("QQBB")
        mpos, mrpos, mq = r.reals(3), r.reals(3), r.reals(4)
        mmat, = r.get("Q")
        mrad, = r.reals(1)
        members.append({"sid": msid, "uid": muid, "visible": bool(mvis), "fixed": bool(mfixed), "pos": mpos,
                        "rpos": mrpos, "q": mq, "material": mmat, "radius": mrad})
    return {"type": UNION_TYPE, "sid": sid, "uid": uid, "visible": bool(vis), "fixed": bool(fixed), "pos": pos,
            "q": q, "v": v, "w": w, "mass": mass, "I": I, "aabb": aabb, "members": members}


def load_body_file(world, filename):
    """Creates the file's bodies in an empty world whose materials exist, as pe's
    BodyBinaryReader::readFile does (system IDs, orientations and velocities from the file;
    masses and inertias from the shapes and materials, as the instantiate* functions do; a
    union's from the file, Union.cpp:100-133)."""
    info = read_body_file(filename)
    B = info["bodies"]
    n = len(B)
    mem = {"owner": [], "sid": [], "uid": [], "mat": [], "radius": [], "rpos": [], "gpos": []}
    union_mass = {}
    dens = np.array([m.density for m in world.materials])
    a = {"shape": np.zeros(n, np.uint8), "mat": np.zeros(n, np.uint32), "fixed": np.zeros(n, np.uint8),
         "param": np.zeros((n, 3)), "pos": np.zeros((n, 3)), "q": np.zeros((n, 4)), "v": np.zeros((n, 3)),
         "w": np.zeros((n, 3)), "inv_mass": np.zeros(n), "I": np.zeros((n, 9)), "Iinv": np.zeros((n, 9)),
         "normal": np.zeros((n, 3)), "d": np.zeros(n), "uid": np.zeros(n, np.uint64)}
    sids = np.zeros(n, np.uint64)
    for i, b in enumerate(B):
        t = b["type"]
        if t == UNION_TYPE:
            if b["fixed"]:
                raise ValueError("union %d is fixed: fixed unions are not supported" % b["sid"])
            if not b["members"]:
                raise ValueError("union %d has no members" % b["sid"])
            for m in b["members"]:
                if m["material"] >= len(dens):
                    raise ValueError("body %d uses material %d, which the world does not have"
                                     % (m["sid"], m["material"]))
                mem["owner"].append(i)
                mem["sid"].append(m["sid"])
                mem["uid"].append(m["uid"])
                mem["mat"].append(m["material"])
                mem["radius"].append(m["radius"])
                mem["rpos"].append(m["rpos"])
                mem["gpos"].append(m["pos"])
            a["shape"][i] = UNION
            a["mat"][i] = b["members"][0]["material"]
            a["pos"][i], a["q"][i], a["v"][i], a["w"][i] = b["pos"], b["q"], b["v"], b["w"]
            a["uid"][i] = b["uid"]
            sids[i] = b["sid"]
            a["I"][i] = b["I"]
            a["inv_mass"][i] = 1.0 / b["mass"]
            union_mass[i] = b["mass"]
            continue
        if b["material"] >= len(dens):
            raise ValueError("body %d uses material %d, which the world does not have" % (b["sid"], b["material"]))
        a["shape"][i] = _SHAPE_OF[t]
        a["mat"][i] = b["material"]
        a["pos"][i], a["q"][i], a["v"][i], a["w"][i] = b["pos"], b["q"], b["v"], b["w"]
        a["uid"][i] = b["uid"]
        sids[i] = b["sid"]
        rho = dens[b["material"]]
        if t == SPHERE_TYPE:
            r = b["param"][0]
            m = 4.0 / 3.0 * math.pi * r * r * r * rho            # SphereBase.h:169-172
            I = 0.4 * m * r * r                                    # SphereBase.h:230-234
            a["param"][i, 0] = r
            a["I"][i, [0, 4, 8]] = I
            a["inv_mass"][i] = 1.0 / m
            a["fixed"][i] = 1 if b["fixed"] else 0
        elif t == BOX_TYPE:
            L = b["param"]
            m = L[0] * L[1] * L[2] * rho                           # BoxBase.h:200-203
            a["param"][i] = L
            a["I"][i, 0] = m / 12.0 * (L[1] * L[1] + L[2] * L[2])  # BoxBase.cpp:152-158
            a["I"][i, 4] = m / 12.0 * (L[0] * L[0] + L[2] * L[2])
            a["I"][i, 8] = m / 12.0 * (L[0] * L[0] + L[1] * L[1])
            a["inv_mass"][i] = 1.0 / m
            a["fixed"][i] = 1 if b["fixed"] else 0
        elif t == CAPSULE_TYPE:
            r, L = b["param"]
            m = r * r * math.pi * (4.0 / 3.0 * r + L) * rho        # CapsuleBase.cpp:60-92
            dn = m / (math.pi * r * r * (4.0 / 3.0 * r + L))       # CapsuleBase.h:196-199
            ms = 4.0 / 3.0 * math.pi * r * r * r * dn              # CapsuleBase.cpp:147-168
            mc = math.pi * r * r * L * dn
            a["param"][i, :2] = (r, L)
            a["I"][i, 0] = r * r * (0.5 * mc + 0.4 * ms)
            a["I"][i, 4] = a["I"][i, 8] = (mc * (0.25 * r * r + 1.0 / 12.0 * L * L) +
                                           ms * (0.4 * r * r + 0.375 * r * L + 0.25 * L * L))
            a["inv_mass"][i] = 1.0 / m
            a["fixed"][i] = 1 if b["fixed"] else 0
        else:                                                      # Plane.cpp:121-142
            nrm = quat_rotate(b["q"], (0.0, 0.0, 1.0))
            g = b["pos"]
            a["normal"][i] = nrm
            a["d"][i] = nrm[0] * g[0] + nrm[1] * g[1] + nrm[2] * g[2]
            a["fixed"][i] = 1
    nonplanar = a["shape"] != PLANE
    a["Iinv"][nonplanar] = mat3_inverse(a["I"][nonplanar])
    # a fixed finite body has no inverse mass or inertia and does not move (RigidBody.cpp:402-418)
    fx = nonplanar & (a["fixed"] == 1)
    a["inv_mass"][fx] = 0.0
    a["Iinv"][fx] = 0.0
    a["v"][fx] = 0.0
    a["w"][fx] = 0.0
    top = max([int(sids.max()) if n else 0] + [int(x) for x in mem["sid"]])
    members = None
    if mem["owner"]:
        members = {k: np.asarray(v) for k, v in mem.items()}
    world.add_bodies(a, sids=sids, next_sid=max(int(info["counter"]), top + 1 if n else 0), members=members,
                     union_mass=union_mass)
    return info
