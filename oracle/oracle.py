"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

The oracle (``oracle/pe_oracle.cpp``) restates the pe reference's HCSITS time
step on the CPU.  Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` import this module, always as the checker.  The product
package ``pe_amd`` never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

SPHERE, BOX, PLANE, CAPSULE, UNION = 1, 2, 3, 4, 5
INELASTIC_FRICTIONLESS = 0
APPROX_INELASTIC_COULOMB_DECOUPLING = 1
APPROX_INELASTIC_COULOMB_ORTHO = 2
INELASTIC_COULOMB_DECOUPLING = 3
INELASTIC_COULOMB_ORTHO = 4
INELASTIC_GMD = 5

_D = ctypes.POINTER(ctypes.c_double)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_I32 = ctypes.POINTER(ctypes.c_int)
_U64 = ctypes.POINTER(ctypes.c_uint64)


class OrcStats(ctypes.Structure):
    _fields_ = [("num_contacts", ctypes.c_size_t), ("max_penetration", ctypes.c_double),
                ("iterations", ctypes.c_size_t), ("last_delta", ctypes.c_double),
                ("t_detection", ctypes.c_double), ("t_contact_caching", ctypes.c_double),
                ("t_body_caching", ctypes.c_double), ("t_solving", ctypes.c_double),
                ("t_integration", ctypes.c_double)]


def build():
    """Compiles the oracle with its Makefile (g++, no FMA, no fast-math)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_libs = {}


def lib(path=None):
    """The oracle library (oracle/_build/liboracle.so), or a variant built from the same source
    at `path` (the tests' planted-mistake builds)."""
    path = path or _LIB_PATH
    if path not in _libs:
        if path == _LIB_PATH and not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        sig = {
            "orc_create": (vp, []), "orc_destroy": (None, [vp]),
            "orc_set_gravity": (None, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "orc_set_damping": (None, [vp, ctypes.c_double]),
            "orc_set_max_iterations": (None, [vp, ctypes.c_size_t]),
            "orc_set_relaxation_model": (None, [vp, ctypes.c_int]),
            "orc_set_erp": (None, [vp, ctypes.c_double]),
            "orc_set_abort_threshold": (None, [vp, ctypes.c_double]),
            "orc_set_trig_mode": (None, [vp, ctypes.c_int]),
            "orc_trig": (ctypes.c_double, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
            "orc_create_material": (ctypes.c_int, [vp] + [ctypes.c_double] * 4),
            "orc_set_pair_friction": (None, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
            "orc_create_sphere": (ctypes.c_int, [vp] + [ctypes.c_double] * 4 + [ctypes.c_int]),
            "orc_create_box": (ctypes.c_int, [vp] + [ctypes.c_double] * 6 + [ctypes.c_int]),
            "orc_create_capsule": (ctypes.c_int, [vp] + [ctypes.c_double] * 5 + [ctypes.c_int]),
            "orc_create_plane": (ctypes.c_int, [vp] + [ctypes.c_double] * 4 + [ctypes.c_int]),
            "orc_create_union": (ctypes.c_int, [vp] + [ctypes.c_double] * 3 + [ctypes.c_int, _D, _D, _I32]),
            "orc_num_members": (ctypes.c_size_t, [vp]),
            "orc_get_members": (None, [vp, _U32, _U64, _I32, _D, _D, _D]),
            "orc_set_members": (None, [vp, _D]),
            "orc_num_bodies": (ctypes.c_size_t, [vp]),
            "orc_get_state": (None, [vp, _D, _D, _D, _D, _D, _D]),
            "orc_set_state": (None, [vp, _D, _D, _D, _D, _D, _D]),
            "orc_get_body_info": (None, [vp, _I32, _U64, _I32, _I32, _D, _D, _D]),
            "orc_set_forces": (None, [vp, _D, _D]),
            "orc_get_aabb": (None, [vp, _D]),
            "orc_step": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_int]),
            "orc_num_contacts": (ctypes.c_size_t, [vp]),
            "orc_set_sweep_order": (None, [vp, ctypes.c_size_t, _U32, _U32]),
            "orc_set_sweep_order_ranked": (None, [vp, ctypes.c_size_t, _U32, _U32, _U32]),
            "orc_get_contacts": (None, [vp, _U32, _U32, _D, _D, _D, _D]),
            "orc_get_impulses": (None, [vp, _D]),
            "orc_detect": (ctypes.c_size_t, [vp, ctypes.c_int]),
            "orc_candidate_pairs": (ctypes.c_size_t, [vp, _U32, _U32, ctypes.c_size_t]),
            "orc_get_stats": (None, [vp, ctypes.POINTER(OrcStats)]),
            "orc_quat_from_axis_angle": (None, [_D, ctypes.c_double, _D]),
            "orc_quat_to_rotation": (None, [_D, _D]),
            "orc_quat_mul": (None, [_D, _D, _D]),
            "orc_quat_rotate": (None, [_D, _D, _D]),
            "orc_mat3_inverse": (None, [_D, _D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _libs[path] = L
    return _libs[path]


def _p(a, typ=_D):
    return a.ctypes.data_as(typ)


class OracleWorld:
    """One pe world restated on the CPU (HashGrids + MaxContacts + HCSITS)."""

    def __init__(self, gravity=(0.0, 0.0, -9.81), max_iterations=100,
                 model=INELASTIC_COULOMB_DECOUPLING, erp=0.7, abort_threshold=0.0, cr_trig=False, lib_path=None):
        self.L = lib(lib_path)
        self.h = ctypes.c_void_p(self.L.orc_create())
        self.L.orc_set_gravity(self.h, *gravity)
        self.L.orc_set_max_iterations(self.h, max_iterations)
        self.L.orc_set_relaxation_model(self.h, model)
        self.L.orc_set_erp(self.h, erp)
        self.L.orc_set_abort_threshold(self.h, abort_threshold)
        self.L.orc_set_trig_mode(self.h, 1 if cr_trig else 0)

    def __del__(self):
        if getattr(self, "h", None) is not None:
            self.L.orc_destroy(self.h)
            self.h = None

    # --- scene -------------------------------------------------------------------------
    def create_material(self, density, cor, csf, cdf):
        return self.L.orc_create_material(self.h, density, cor, csf, cdf)

    def set_pair_friction(self, m1, m2, csf, cdf):
        self.L.orc_set_pair_friction(self.h, m1, m2, csf, cdf)

    def create_sphere(self, pos, r, mat):
        return self.L.orc_create_sphere(self.h, float(pos[0]), float(pos[1]), float(pos[2]), float(r), mat)

    def create_box(self, pos, lengths, mat):
        return self.L.orc_create_box(self.h, *[float(x) for x in pos], *[float(x) for x in lengths], mat)

    def create_capsule(self, pos, r, length, mat):
        """A capsule whose axis is the body x-axis (CapsuleBase.cpp:60-92)."""
        return self.L.orc_create_capsule(self.h, *[float(x) for x in pos], float(r), float(length), mat)

    def create_plane(self, normal, d, mat):
        return self.L.orc_create_plane(self.h, *[float(x) for x in normal], float(d), mat)

    def create_union(self, pos, disp, radius, mat):
        """A union of spheres at offsets disp (k x 3) whose centre of mass goes to pos."""
        d = np.ascontiguousarray(disp, np.float64).reshape(-1, 3)
        r = np.ascontiguousarray(radius, np.float64).reshape(-1)
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(mat, np.int32), r.shape))
        idx = self.L.orc_create_union(self.h, *[float(x) for x in pos], len(r), _p(d), _p(r), _p(m, _I32))
        if idx < 0:
            raise ValueError("a union needs at least one sphere")
        return idx

    def members(self):
        """Union members: owner body, system ID, material, radius, relative and global position."""
        k = int(self.L.orc_num_members(self.h))
        m = {"owner": np.zeros(k, np.uint32), "sid": np.zeros(k, np.uint64), "mat": np.zeros(k, np.int32),
             "radius": np.zeros(k), "rpos": np.zeros((k, 3)), "gpos": np.zeros((k, 3))}
        if k:
            self.L.orc_get_members(self.h, _p(m["owner"], _U32), _p(m["sid"], _U64), _p(m["mat"], _I32),
                                   _p(m["radius"]), _p(m["rpos"]), _p(m["gpos"]))
        return m

    def set_members(self, gpos):
        g = np.ascontiguousarray(gpos, np.float64)
        self.L.orc_set_members(self.h, _p(g))

    @property
    def n(self):
        return int(self.L.orc_num_bodies(self.h))

    # --- state -------------------------------------------------------------------------
    def get_state(self):
        """Dynamic state; with unions also "mgpos", the members' global positions."""
        n = self.n
        st = {"pos": np.zeros((n, 3)), "q": np.zeros((n, 4)), "v": np.zeros((n, 3)),
              "w": np.zeros((n, 3)), "normal": np.zeros((n, 3)), "d": np.zeros(n)}
        self.L.orc_get_state(self.h, _p(st["pos"]), _p(st["q"]), _p(st["v"]), _p(st["w"]),
                             _p(st["normal"]), _p(st["d"]))
        if self.L.orc_num_members(self.h):
            st["mgpos"] = self.members()["gpos"]
        return st

    def set_state(self, st):
        arrs = [np.ascontiguousarray(st[k], dtype=np.float64) for k in ("pos", "q", "v", "w", "normal", "d")]
        self.L.orc_set_state(self.h, *[_p(a) for a in arrs])
        if "mgpos" in st and self.L.orc_num_members(self.h):
            self.set_members(st["mgpos"])

    def set_forces(self, force, torque):
        """Force and torque per body (world frame) for the next step only."""
        f = np.ascontiguousarray(force, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(torque, np.float64).reshape(-1, 3)
        assert f.shape[0] == self.n and t.shape[0] == self.n
        self.L.orc_set_forces(self.h, _p(f), _p(t))

    def body_info(self):
        n = self.n
        info = {"shape": np.zeros(n, np.int32), "sid": np.zeros(n, np.uint64), "mat": np.zeros(n, np.int32),
                "fixed": np.zeros(n, np.int32), "param": np.zeros((n, 3)), "inv_mass": np.zeros(n),
                "inv_inertia": np.zeros((n, 9))}
        self.L.orc_get_body_info(self.h, _p(info["shape"], _I32), _p(info["sid"], _U64), _p(info["mat"], _I32),
                                 _p(info["fixed"], _I32), _p(info["param"]), _p(info["inv_mass"]),
                                 _p(info["inv_inertia"]))
        return info

    def aabb(self):
        a = np.zeros((self.n, 6))
        self.L.orc_get_aabb(self.h, _p(a))
        return a

    # --- stepping ----------------------------------------------------------------------
    def step(self, dt, nsteps=1):
        rc = self.L.orc_step(self.h, dt, nsteps)
        if rc == -1:
            raise ValueError("Invalid time step size")
        if rc != 0:
            raise RuntimeError("oracle step failed (unsupported contact pair)")

    def set_sweep_order(self, b1, b2):
        """Relax the next step's contacts in this (body1, body2) order."""
        a = np.ascontiguousarray(b1, np.uint32)
        b = np.ascontiguousarray(b2, np.uint32)
        self.L.orc_set_sweep_order(self.h, len(a), _p(a, _U32), _p(b, _U32))

    def set_sweep_order_ranked(self, b1, b2, rank):
        """Relax the next step's contacts in this order; entry k is the rank[k]-th contact (in
        narrowphase order) of the pair (b1[k], b2[k])."""
        a = np.ascontiguousarray(b1, np.uint32)
        b = np.ascontiguousarray(b2, np.uint32)
        r = np.ascontiguousarray(rank, np.uint32)
        self.L.orc_set_sweep_order_ranked(self.h, len(a), _p(a, _U32), _p(b, _U32), _p(r, _U32))

    def contacts(self):
        nc = int(self.L.orc_num_contacts(self.h))
        c = {"b1": np.zeros(nc, np.uint32), "b2": np.zeros(nc, np.uint32), "pos": np.zeros((nc, 3)),
             "normal": np.zeros((nc, 3)), "dist": np.zeros(nc), "mu": np.zeros(nc)}
        self.L.orc_get_contacts(self.h, _p(c["b1"], _U32), _p(c["b2"], _U32), _p(c["pos"]),
                                _p(c["normal"]), _p(c["dist"]), _p(c["mu"]))
        return c

    def impulses(self):
        nc = int(self.L.orc_num_contacts(self.h))
        p = np.zeros((nc, 3))
        self.L.orc_get_impulses(self.h, _p(p))
        return p

    def detect(self, exhaustive=False):
        n = self.L.orc_detect(self.h, 1 if exhaustive else 0)
        if n == ctypes.c_size_t(-1).value:
            raise RuntimeError("oracle detection failed (unsupported contact pair)")
        return self.contacts()

    def candidate_pairs(self):
        n = int(self.L.orc_candidate_pairs(self.h, None, None, 0))
        a = np.zeros(n, np.uint32)
        b = np.zeros(n, np.uint32)
        self.L.orc_candidate_pairs(self.h, _p(a, _U32), _p(b, _U32), n)
        return a, b

    def stats(self):
        s = OrcStats()
        self.L.orc_get_stats(self.h, ctypes.byref(s))
        return {f: getattr(s, f) for f, _ in OrcStats._fields_}


def trig(fn, x, y=0.0, correctly_rounded=False):
    """sin (0), cos (1), atan2(y, x) (2), acos (3): C library or correctly rounded."""
    return lib().orc_trig(int(fn), float(x), float(y), 1 if correctly_rounded else 0)


def quat_from_axis_angle(axis, angle):
    a = np.ascontiguousarray(axis, dtype=np.float64)
    q = np.zeros(4)
    lib().orc_quat_from_axis_angle(_p(a), float(angle), _p(q))
    return q


def quat_to_rotation(q):
    qa = np.ascontiguousarray(q, dtype=np.float64)
    R = np.zeros(9)
    lib().orc_quat_to_rotation(_p(qa), _p(R))
    return R.reshape(3, 3)


def quat_mul(a, b):
    x = np.ascontiguousarray(a, dtype=np.float64)
    y = np.ascontiguousarray(b, dtype=np.float64)
    o = np.zeros(4)
    lib().orc_quat_mul(_p(x), _p(y), _p(o))
    return o


def quat_rotate(q, v):
    x = np.ascontiguousarray(q, dtype=np.float64)
    y = np.ascontiguousarray(v, dtype=np.float64)
    o = np.zeros(3)
    lib().orc_quat_rotate(_p(x), _p(y), _p(o))
    return o


def mat3_inverse(m):
    x = np.ascontiguousarray(m, dtype=np.float64).reshape(9)
    o = np.zeros(9)
    lib().orc_mat3_inverse(_p(x), _p(o))
    return o.reshape(3, 3)
