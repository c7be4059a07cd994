"""Host-side mirror of pe's World / CollisionSystem API over the HIP library.

Names, argument meanings and error behaviour follow the reference:
``createMaterial`` (src/core/Materials.cpp:175-247), ``createSphere``
(src/core/rigidbody/Sphere.cpp:1187), ``createBox`` (src/core/rigidbody/Box.cpp:1312),
``createPlane`` (src/core/rigidbody/Plane.cpp:1202-1240), ``World::run`` /
``World::simulationStep`` (src/core/World.cpp:330-378) and the HCSITS collision
system setters (pe/core/collisionsystem/HardContactSemiImplicitTimesteppingSolvers.h:793-840).
Mass properties are computed on the host exactly as pe computes them; the step
itself always runs on the GPU.
"""
import ctypes
import math

import numpy as np

from . import _abi
from ._abi import BOX, CAPSULE, PLANE, SPHERE, UNION, check, ptr

CONTACT_THRESHOLD = 1e-8   # pe/core/Thresholds.h:151
ACCURACY = 1e-8            # pe/util/Limits.h:385

# CollisionSystem<HCSITS>::RelaxationModel (HCSITS.h:177-184)
InelasticFrictionlessContact = 0
ApproximateInelasticCoulombContactByDecoupling = 1
ApproximateInelasticCoulombContactByOrthogonalProjections = 2
InelasticCoulombContactByDecoupling = 3
InelasticCoulombContactByOrthogonalProjections = 4
InelasticGeneralizedMaximumDissipationContact = 5


def mat3_inverse(m):
    """Cofactor inverse of (..., 9) row-major matrices in pe's evaluation order
    (Matrix3x3::getInverse, pe/math/Matrix3x3.h:1084-1104)."""
    v = [m[..., k] for k in range(9)]
    det = v[0] * ((v[4] * v[8]) - (v[7] * v[5])) + v[1] * ((v[5] * v[6]) - (v[8] * v[3])) + \
        v[2] * ((v[3] * v[7]) - (v[4] * v[6]))
    det = 1.0 / det
    out = np.stack([det * ((v[4] * v[8]) - (v[5] * v[7])), det * ((v[7] * v[2]) - (v[8] * v[1])),
                    det * ((v[1] * v[5]) - (v[2] * v[4])), det * ((v[5] * v[6]) - (v[3] * v[8])),
                    det * ((v[8] * v[0]) - (v[6] * v[2])), det * ((v[2] * v[3]) - (v[0] * v[5])),
                    det * ((v[3] * v[7]) - (v[4] * v[6])), det * ((v[6] * v[1]) - (v[7] * v[0])),
                    det * ((v[0] * v[4]) - (v[1] * v[3]))], axis=-1)
    return out


def granular_particle_spheres(rng, radius, material):
    """The spheres of pe's granular particle (createGranularParticle, Union.cpp:2376-2402): two to
    four spheres of radius U(0.6, 0.8) * radius, each offset from the particle centre along a
    normalised U(-1, 1)^3 direction by radius minus its own radius.  The draws come from rng (a
    numpy Generator), not from pe's Boost generator, so a particle equals pe's only in
    distribution.  Returns [(offset, radius, material), ...] for World.createUnion."""
    if not radius > 0:
        raise ValueError("Invalid sphere radius")
    number = int(rng.integers(2, 5))
    rmin, rmax = 0.6 * radius, 0.8 * radius
    spheres = []
    for _ in range(number):
        rad = float(rng.uniform(rmin, rmax))
        d = rng.uniform(-1.0, 1.0, size=3)
        ilen = 1.0 / math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])   # Vector3::normalize
        d = [float(x) * ilen * (radius - rad) for x in d]                  # Vector3::scale
        spheres.append((d, rad, material))
    return spheres


def _mat3_mul(l, r):
    """Row-major 3x3 product in pe's evaluation order (Matrix3x3.h:2241-2253)."""
    return [l[3 * i] * r[j] + l[3 * i + 1] * r[3 + j] + l[3 * i + 2] * r[6 + j] for i in range(3) for j in range(3)]


def quat_axis_angle(axis, angle):
    """Quaternion(axis, angle), pe/math/Quaternion.h:274-294 (host libm)."""
    if abs(angle) < 1e-15:
        return np.array([1.0, 0.0, 0.0, 0.0])
    sina = math.sin(angle * 0.5)
    cosa = math.cos(angle * 0.5)
    a = normalize(axis)
    return np.array([cosa, sina * a[0], sina * a[1], sina * a[2]])


def normalize(v):
    """Vector3::normalize (pe/math/Vector3.h:994-1010): reciprocal length, then scale."""
    x, y, z = float(v[0]), float(v[1]), float(v[2])
    ln = math.sqrt(x * x + y * y + z * z)
    if ln == 0.0:
        return np.array([x, y, z])
    il = 1.0 / ln
    return np.array([x * il, y * il, z * il])


class Material:
    def __init__(self, name, density, cor, csf, cdf, poisson, young, stiffness, dampingN, dampingT):
        self.name, self.density, self.cor, self.csf, self.cdf = name, density, cor, csf, cdf
        self.poisson, self.young, self.stiffness, self.dampingN, self.dampingT = poisson, young, stiffness, dampingN, dampingT


def _per_body(a, n, k, name, dtype=float):
    """a as a contiguous (n, k) array (n values when k == 0); the C-ABI reads exactly that many
    values, so any other size is refused before it reaches the library."""
    arr = np.ascontiguousarray(a, dtype)
    want = n * k if k else n
    if arr.size != want:
        raise ValueError("%s: expected %d values (%d bodies), got %d" % (name, want, n, arr.size))
    return arr


def union_param(rpos, radii):
    """A union's radii for the broadphase and for the solver's body records: the radius of a
    sphere about its centre of mass that holds every member (widened by 1e-12 so that rounding
    never lets a member poke out) and the smallest member radius."""
    bound = max(math.sqrt(float(rp[0]) * float(rp[0]) + float(rp[1]) * float(rp[1]) + float(rp[2]) * float(rp[2]))
                + float(r) for rp, r in zip(rpos, radii))
    return (bound * (1.0 + 1e-12), min(float(r) for r in radii), 0.0)


class World:
    """One pe world (``theWorld()`` + ``theCollisionSystem()`` of the reference)."""

    def __init__(self, device=0):
        self.device = device
        self.materials = []
        self._csf = np.zeros((0, 0))
        self._cdf = np.zeros((0, 0))
        self._gravity = (0.0, 0.0, 0.0)
        self._damping = 1.0
        self._max_iterations = 100
        self._model = ApproximateInelasticCoulombContactByDecoupling
        self._erp = 0.7
        self._relax_param = 0.9
        self._abort = 0.0
        self._max_sub = 20
        self._sweep_mode = _abi.SWEEP_COLOURED
        self._exact_trig = 0
        self._schedule = 0
        self._next_sid = 1
        self._shape, self._sid, self._mat, self._fixed = [], [], [], []
        self._param, self._pos, self._q, self._v, self._w = [], [], [], [], []
        self._inv_mass, self._I, self._Iinv, self._normal, self._d = [], [], [], [], []
        self._uid = []
        # union members (spheres), grouped by union in creation order
        self._m_owner, self._m_sid, self._m_mat, self._m_rad, self._m_rpos, self._m_gpos = [], [], [], [], [], []
        self._m_uid = []
        self._u_mass = {}        # union body index -> total mass (Union::getMass)
        self._arrays = None      # host SoA snapshot (numpy) when built
        self._ctx = None
        self._bodies_dirty = True
        self._params_dirty = True
        self._state_on_device = False
        self.timestep = 0

    # ------------------------------------------------------------------ settings
    def setGravity(self, gx, gy=None, gz=None):
        if gy is None:
            gx, gy, gz = gx
        self._gravity = (float(gx), float(gy), float(gz))
        self._params_dirty = True

    def getGravity(self):
        return self._gravity

    def setDamping(self, damping):
        if damping < 0 or damping > 1:
            raise ValueError("Invalid damping value")
        self._damping = float(damping)
        self._params_dirty = True

    # collision-system setters (HCSITS.h:770-840)
    def setMaxIterations(self, n):
        self._max_iterations = int(n)
        self._params_dirty = True

    def setRelaxationModel(self, model):
        self._model = int(model)
        self._params_dirty = True

    def setErrorReductionParameter(self, erp):
        if erp < 0 or erp > 1:
            raise ValueError("Error reduction parameter out of range.")
        self._erp = float(erp)
        self._params_dirty = True

    def setRelaxationParameter(self, f):
        if f <= 0:
            raise ValueError("Relaxation parameter must be positive.")
        self._relax_param = float(f)
        self._params_dirty = True

    def setAbortThreshold(self, thr):
        """Convergence stop on the L-inf impulse change of a sweep (the reference's check at
        HCSITS.h:1935-1946, compiled out there; 0 = always run max iterations)."""
        self._abort = float(thr)
        self._params_dirty = True

    def setSweepMode(self, mode):
        self._sweep_mode = int(mode)
        self._params_dirty = True

    def setSweepSchedule(self, schedule):
        """Coloured mode: 0 auto, 1 one launch per colour, 2 dataflow, 3 one workgroup, 4 islands
        (pehip_sweep_schedule)."""
        if schedule not in (0, 1, 2, 3, 4):
            raise ValueError("Invalid sweep schedule.")
        self._schedule = int(schedule)
        self._params_dirty = True

    def setMaxSubIterations(self, n):
        """Golden-section steps of relaxation models 3 and 5 (pe's maxSubIterations_,
        HCSITS.h:305, 423: fixed at 20 there, with no setter)."""
        if int(n) < 0:
            raise ValueError("Invalid number of sub-iterations")
        self._max_sub = int(n)
        self._params_dirty = True

    def setExactTrig(self, on=True):
        """Correctly rounded sin/cos/atan2/acos in the solver (reproduces the host libm)."""
        self._exact_trig = 1 if on else 0
        self._params_dirty = True

    # ------------------------------------------------------------------ materials
    def createMaterial(self, name, density, cor, csf, cdf, poisson=0.5, young=1.0, stiffness=1.0,
                       dampingN=0.0, dampingT=0.0):
        """createMaterial, Materials.cpp:175-247 (pair friction = csf_i + csf_j)."""
        if any(m.name == name for m in self.materials):
            raise ValueError("Material of that name already exists!")
        if density <= 0:
            raise ValueError("Invalid material density!")
        if cor < 0 or cor > 1:
            raise ValueError("Invalid coefficient of restitution!")
        if csf < 0:
            raise ValueError("Invalid coefficient of static friction!")
        if cdf < 0:
            raise ValueError("Invalid coefficient of dynamic friction!")
        if poisson < -1 or poisson > 0.5:
            raise ValueError("Invalid Poisson's ratio")
        if young <= 0:
            raise ValueError("Invalid Young's modulus")
        if stiffness <= 0:
            raise ValueError("Invalid stiffness")
        if dampingN < 0 or dampingT < 0:
            raise ValueError("Invalid damping coefficients")
        self.materials.append(Material(name, float(density), float(cor), float(csf), float(cdf), poisson, young,
                                       stiffness, dampingN, dampingT))
        n = len(self.materials)
        s = np.zeros((n, n))
        d = np.zeros((n, n))
        s[:n - 1, :n - 1] = self._csf
        d[:n - 1, :n - 1] = self._cdf
        k = n - 1
        s[k, k] = csf + csf
        d[k, k] = cdf + cdf
        for i in range(k):
            s[k, i] = s[i, k] = self.materials[i].csf + csf
            d[k, i] = d[i, k] = self.materials[i].cdf + cdf
        self._csf, self._cdf = s, d
        self._bodies_dirty = True
        return k

    def setPairFriction(self, m1, m2, csf, cdf):
        """Explicit pair coefficients (Materials.h:746)."""
        self._csf[m1, m2] = self._csf[m2, m1] = csf
        self._cdf[m1, m2] = self._cdf[m2, m1] = cdf
        self._bodies_dirty = True

    # ------------------------------------------------------------------ bodies
    def _sync_from_device(self):
        if self._state_on_device and self._ctx is not None:
            st = self.getState()
            mg = st.pop("mgpos", None)
            self._arrays.update(st)
            if mg is not None:
                self._m_gpos = list(mg)
            self._state_on_device = False

    def _materialise_lists(self):
        if self._arrays is not None:
            self._sync_from_device()
            a = self._arrays
            self._shape, self._sid, self._mat, self._fixed = list(a["shape"]), list(a["sid"]), list(a["mat"]), list(a["fixed"])
            self._param, self._pos, self._q = list(a["param"]), list(a["pos"]), list(a["q"])
            self._v, self._w, self._inv_mass = list(a["v"]), list(a["w"]), list(a["inv_mass"])
            self._I, self._Iinv, self._normal, self._d = list(a["I"]), list(a["Iinv"]), list(a["normal"]), list(a["d"])
            self._uid = list(a["uid"]) if "uid" in a else [0] * len(self._shape)
            self._arrays = None

    def _add(self, shape, mat, fixed, param, pos, q, inv_mass, I, Iinv, normal, d, uid=0):
        self._materialise_lists()
        self._uid.append(int(uid))
        sid = self._next_sid
        self._next_sid += 1
        self._shape.append(shape); self._sid.append(sid); self._mat.append(mat); self._fixed.append(1 if fixed else 0)
        self._param.append(np.asarray(param, float)); self._pos.append(np.asarray(pos, float))
        self._q.append(np.asarray(q, float)); self._v.append(np.zeros(3)); self._w.append(np.zeros(3))
        self._inv_mass.append(float(inv_mass)); self._I.append(np.asarray(I, float)); self._Iinv.append(np.asarray(Iinv, float))
        self._normal.append(np.asarray(normal, float)); self._d.append(float(d))
        self._bodies_dirty = True
        return len(self._shape) - 1

    def createSphere(self, uid, gpos, radius, material):
        """createSphere (Sphere.cpp:1187) with SphereBase's mass properties (SphereBase.cpp:64-87)."""
        if radius <= 0:
            raise ValueError("Invalid sphere radius")
        r = float(radius)
        m = 4.0 / 3.0 * math.pi * r * r * r * self.materials[material].density
        I = np.zeros(9)
        I[0] = I[4] = I[8] = 0.4 * m * r * r
        return self._add(SPHERE, material, False, (r, 0.0, 0.0), gpos, (1.0, 0.0, 0.0, 0.0), 1.0 / m, I,
                         mat3_inverse(I), (0.0, 0.0, 0.0), 0.0, uid)

    def createBox(self, uid, gpos, lengths, material):
        """createBox (Box.cpp:1312) with BoxBase's mass properties (BoxBase.cpp:60-160)."""
        lx, ly, lz = (float(x) for x in lengths)
        if lx <= 0 or ly <= 0 or lz <= 0:
            raise ValueError("Invalid side length")
        m = lx * ly * lz * self.materials[material].density
        I = np.zeros(9)
        I[0] = m / 12.0 * (ly * ly + lz * lz)
        I[4] = m / 12.0 * (lx * lx + lz * lz)
        I[8] = m / 12.0 * (lx * lx + ly * ly)
        return self._add(BOX, material, False, (lx, ly, lz), gpos, (1.0, 0.0, 0.0, 0.0), 1.0 / m, I,
                         mat3_inverse(I), (0.0, 0.0, 0.0), 0.0, uid)

    def createCapsule(self, uid, gpos, radius, length, material):
        """createCapsule (Capsule.cpp:1219) with CapsuleBase's mass properties (CapsuleBase.cpp:60-92,
        147-168): the axis is the body x-axis; the moments combine the cylinder and the end caps,
        with the density recomputed from the mass as the reference does."""
        r, L = float(radius), float(length)
        if r <= 0:
            raise ValueError("Invalid capsule radius")
        if L <= 0:
            raise ValueError("Invalid capsule length")
        m = r * r * math.pi * (4.0 / 3.0 * r + L) * self.materials[material].density
        density = m / (math.pi * r * r * (4.0 / 3.0 * r + L))
        sphere_mass = 4.0 / 3.0 * math.pi * r * r * r * density
        cylinder_mass = math.pi * r * r * L * density
        I = np.zeros(9)
        I[0] = r * r * (0.5 * cylinder_mass + 0.4 * sphere_mass)
        I[4] = I[8] = (cylinder_mass * (0.25 * r * r + 1.0 / 12.0 * L * L) +
                       sphere_mass * (0.4 * r * r + 0.375 * r * L + 0.25 * L * L))
        return self._add(CAPSULE, material, False, (r, L, 0.0), gpos, (1.0, 0.0, 0.0, 0.0), 1.0 / m, I,
                         mat3_inverse(I), (0.0, 0.0, 0.0), 0.0, uid)

    def createUnion(self, uid, gpos, spheres):
        """A union of spheres, as pe builds one inside pe_CREATE_UNION and places it with
        Union::setPosition (Union.cpp:251-272, 1482-1545; UnionBase.cpp:240-360).

        spheres: (offset, radius, material) or (offset, radius, material, uid) per sphere, offsets
        before placement; a sphere without a user ID takes the union's (as createGranularParticle's
        spheres do, Union.cpp:2376-2402).  The union takes the next system ID, then each sphere its own.  After the last sphere: total mass and centre
        of mass, the spheres' positions relative to it, the moment of inertia about it (parallel
        axes), all with the identity orientation; then the centre of mass moves to gpos and every
        sphere by the same displacement."""
        muid = [int(t[3]) if len(t) > 3 else int(uid) for t in spheres]
        spheres = [(np.asarray(t[0], float).reshape(3), float(t[1]), int(t[2])) for t in spheres]
        if not spheres:
            raise ValueError("Invalid number of spheres")
        for _, r, m in spheres:
            if not r > 0:
                raise ValueError("Invalid sphere radius")
            if not 0 <= m < len(self.materials):
                raise ValueError("Invalid material ID")
        disp = [o for o, _, _ in spheres]
        mass_k, inert_k = [], []
        for _, r, m in spheres:
            mk = 4.0 / 3.0 * math.pi * r * r * r * self.materials[m].density
            mass_k.append(mk)
            inert_k.append(0.4 * mk * r * r)
        n = len(spheres)
        if n == 1:
            com = [float(disp[0][0]), float(disp[0][1]), float(disp[0][2])]
            mass = mass_k[0]
        else:
            com, mass = [0.0, 0.0, 0.0], 0.0
            for k in range(n):
                mass += mass_k[k]
                for j in range(3):
                    com[j] += float(disp[k][j]) * mass_k[k]
            tmp = 1.0 / mass          # Vector3::operator/=, Vector3.h:836-848
            com = [c * tmp for c in com]
        inv_mass = 1.0 / mass
        eye = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
        rpos = []
        for k in range(n):
            d = [float(disp[k][j]) - com[j] for j in range(3)]
            rpos.append([eye[3 * i] * d[0] + eye[3 * i + 1] * d[1] + eye[3 * i + 2] * d[2] for i in range(3)])
        I = [0.0] * 9
        for k in range(n):
            mk = mass_k[k]
            pos = [float(disp[k][j]) - com[j] for j in range(3)]
            I[0] += inert_k[k]; I[4] += inert_k[k]; I[8] += inert_k[k]
            I[0] += mk * (pos[1] * pos[1] + pos[2] * pos[2])
            I[1] -= mk * pos[0] * pos[1]
            I[2] -= mk * pos[0] * pos[2]
            I[3] -= mk * pos[0] * pos[1]
            I[4] += mk * (pos[0] * pos[0] + pos[2] * pos[2])
            I[5] -= mk * pos[1] * pos[2]
            I[6] -= mk * pos[0] * pos[2]
            I[7] -= mk * pos[1] * pos[2]
            I[8] += mk * (pos[0] * pos[0] + pos[1] * pos[1])
        I = _mat3_mul(_mat3_mul(eye, I), eye)   # trans(R) * I * R with R = identity
        Iinv = mat3_inverse(np.asarray(I))
        g = [float(x) for x in np.asarray(gpos, float).reshape(3)]
        dp = [g[j] - com[j] for j in range(3)]
        idx = self._add(UNION, spheres[0][2], False, union_param(rpos, [r for _, r, _ in spheres]), g,
                        (1.0, 0.0, 0.0, 0.0), inv_mass, np.asarray(I), Iinv, (0.0, 0.0, 0.0), 0.0, uid)
        self._u_mass[idx] = mass
        for k in range(n):
            self._m_owner.append(idx)
            self._m_sid.append(self._next_sid)
            self._next_sid += 1
            self._m_uid.append(muid[k])
            self._m_mat.append(spheres[k][2])
            self._m_rad.append(spheres[k][1])
            self._m_rpos.append(np.asarray(rpos[k], float))
            self._m_gpos.append(np.asarray([float(disp[k][j]) + dp[j] for j in range(3)]))
        return idx

    def createGranularParticle(self, uid, gpos, radius, material, rng=None):
        """createGranularParticle (Union.cpp:2376-2402); see granular_particle_spheres."""
        rng = np.random.default_rng() if rng is None else rng
        return self.createUnion(uid, gpos, granular_particle_spheres(rng, radius, material))

    def members(self):
        """Union members: owner body, system ID, user ID, material, radius, relative and global
        position."""
        self._sync_from_device()
        k = len(self._m_owner)
        return {"owner": np.asarray(self._m_owner, np.uint32), "sid": np.asarray(self._m_sid, np.uint64),
                "uid": np.asarray(self._m_uid, np.uint64),
                "mat": np.asarray(self._m_mat, np.uint32), "radius": np.asarray(self._m_rad, float),
                "rpos": np.asarray(self._m_rpos, float).reshape(k, 3), "gpos": np.asarray(self._m_gpos, float).reshape(k, 3)}

    def createPlane(self, uid, normal, d, material):
        """createPlane (Plane.cpp:1202-1240) and the Plane ctor (Plane.cpp:77-102)."""
        nv = np.asarray(normal, float)
        if float(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]) == 0.0:
            raise ValueError("Invalid plane normal!")
        n = normalize(nv)
        d = float(d)
        if n[0] * n[0] + n[1] * n[1] < ACCURACY:
            q = quat_axis_angle((1.0, 0.0, 0.0), math.acos(n[2]))
        else:
            q = quat_axis_angle((-n[1], n[0], 0.0), math.acos(n[2]))
        return self._add(PLANE, material, True, (0.0, 0.0, 0.0), n * d, q, 0.0, np.zeros(9), np.zeros(9), n, d, uid)

    def setFixed(self, body, fixed=True):
        """RigidBody::fix (RigidBody.cpp:402-418): zero inverse mass and inertia, zero velocity."""
        self._materialise_lists()
        if self._shape[body] == PLANE or not fixed:
            return
        if self._shape[body] == UNION:
            raise ValueError("fixed unions are not supported")
        self._fixed[body] = 1
        self._inv_mass[body] = 0.0
        self._Iinv[body] = np.zeros(9)
        self._v[body] = np.zeros(3)
        self._w[body] = np.zeros(3)
        self._bodies_dirty = True

    def destroy(self, body):
        """destroy(BodyID) (RigidBody.cpp:766-775) -> World::remove -> HCSITS::remove (World.cpp:298-306,
        HCSITS.h:1346-1373): the body leaves the world.  pe's BodyStorage::remove
        (BodyStorage.h:563-573) moves the last body into the removed body's place, so the last body
        takes index ``body``.  A union leaves with its members.  The next step uploads the
        remaining bodies."""
        self._materialise_lists()
        n = len(self._shape)
        body = int(body)
        if not 0 <= body < n:
            raise IndexError("invalid body index %d (%d bodies)" % (body, n))
        last = n - 1
        for lst in (self._shape, self._sid, self._mat, self._fixed, self._param, self._pos, self._q, self._v,
                    self._w, self._inv_mass, self._I, self._Iinv, self._normal, self._d, self._uid):
            lst[body] = lst[last]
            lst.pop()
        if self._m_owner:
            # the removed union's members leave; the moved body's follow it; members stay grouped
            # by union in body order, each union's in its own order (a stable sort)
            owner = [body if o == last else o for o in self._m_owner]
            keep = sorted((k for k, o in enumerate(self._m_owner) if o != body), key=lambda k: owner[k])
            self._m_owner = [owner[k] for k in keep]
            for name in ("_m_sid", "_m_uid", "_m_mat", "_m_rad", "_m_rpos", "_m_gpos"):
                src = getattr(self, name)
                setattr(self, name, [src[k] for k in keep])
        self._u_mass.pop(body, None)
        if last != body and last in self._u_mass:
            self._u_mass[body] = self._u_mass.pop(last)
        self._bodies_dirty = True

    def unionMass(self, body):
        """Total mass of a union (Union::getMass), as created or as read from a body file."""
        return self._u_mass[body]

    def add_bodies(self, arrays, sids=None, next_sid=None, members=None, union_mass=None):
        """Bulk creation from SoA arrays in body order (as a sequence of create* calls; or with
        given system IDs, as a body file reader instantiates bodies).

        Unions come with their stored mass properties, as pe's reader instantiates them
        (Union.cpp:100-133, Union::add skipping the recomputation inside pe_INSTANTIATE_UNION,
        Union.cpp:1522-1524): a union row's inv_mass, I and Iinv are taken as given, union_mass
        maps its body index to its total mass, and members (a dict as ``members()`` returns, grouped
        by union in body order) holds its spheres.  A union's bounding radii are its members'."""
        self._materialise_lists()
        base = len(self._shape)
        n = len(arrays["shape"])
        if base:
            raise ValueError("bulk creation only into an empty world")
        # the device upload reads a fixed number of values per body from each array
        for key, k in (("mat", 1), ("fixed", 1), ("param", 3), ("pos", 3), ("q", 4), ("v", 3), ("w", 3),
                       ("inv_mass", 1), ("I", 9), ("Iinv", 9), ("normal", 3), ("d", 1)):
            if key not in arrays or np.asarray(arrays[key]).size != k * n:
                raise ValueError("body array %s: expected %d values (%d bodies)" % (key, k * n, n))
        if sids is not None and np.asarray(sids).size != n:
            raise ValueError("sids: expected %d values" % n)
        shape = np.asarray(arrays["shape"])
        unions = np.flatnonzero(shape == UNION)
        arrays = dict(arrays)
        if len(unions):
            if members is None:
                raise ValueError("unions need their members")
            owner = np.asarray(members["owner"], np.int64)
            if np.any(np.diff(owner) < 0) or np.any(shape[owner] != UNION):
                raise ValueError("members must be grouped by union in body order")
            if set(owner.tolist()) != set(unions.tolist()):
                raise ValueError("union without members")
            if np.any(np.asarray(arrays["fixed"])[unions]):
                raise ValueError("fixed unions are not supported")
            k = len(owner)
            for key in ("sid", "mat", "radius"):
                if np.asarray(members[key]).size != k:
                    raise ValueError("member array %s: expected %d values (%d members)" % (key, k, k))
            rad = np.asarray(members["radius"], float)
            rpos = np.asarray(members["rpos"], float).reshape(k, 3)
            arrays["param"] = np.array(arrays["param"], float)
            for u in unions:
                sel = np.flatnonzero(owner == u)
                arrays["param"][u] = union_param(rpos[sel], rad[sel])
            self._m_owner = owner.tolist()
            self._m_sid = [int(x) for x in members["sid"]]
            self._m_uid = [int(x) for x in members.get("uid", np.asarray(arrays["uid"])[owner])]
            self._m_mat = [int(x) for x in members["mat"]]
            self._m_rad = rad.tolist()
            self._m_rpos = list(rpos)
            self._m_gpos = list(np.asarray(members["gpos"], float).reshape(k, 3))
            im = np.asarray(arrays["inv_mass"], float)
            self._u_mass = {int(u): (float(union_mass[int(u)]) if union_mass is not None else 1.0 / float(im[u]))
                            for u in unions}
        self._arrays = arrays
        if sids is None:
            self._arrays["sid"] = np.arange(self._next_sid, self._next_sid + n, dtype=np.uint64)
            self._next_sid += n
        else:
            self._arrays["sid"] = np.asarray(sids, np.uint64).copy()
            self._next_sid = int(next_sid) if next_sid is not None else int(self._arrays["sid"].max()) + 1
        self._bodies_dirty = True
        return np.arange(n)

    @property
    def size(self):
        return len(self._arrays["shape"]) if self._arrays is not None else len(self._shape)

    def host_arrays(self):
        """Body arrays (numpy SoA) in body order."""
        if self._arrays is None:
            n = len(self._shape)
            self._arrays = {
                "shape": np.asarray(self._shape, np.uint8), "sid": np.asarray(self._sid, np.uint64),
                "mat": np.asarray(self._mat, np.uint32), "fixed": np.asarray(self._fixed, np.uint8),
                "param": np.asarray(self._param, float).reshape(n, 3), "pos": np.asarray(self._pos, float).reshape(n, 3),
                "q": np.asarray(self._q, float).reshape(n, 4), "v": np.asarray(self._v, float).reshape(n, 3),
                "w": np.asarray(self._w, float).reshape(n, 3), "inv_mass": np.asarray(self._inv_mass, float),
                "I": np.asarray(self._I, float).reshape(n, 9), "Iinv": np.asarray(self._Iinv, float).reshape(n, 9),
                "normal": np.asarray(self._normal, float).reshape(n, 3), "d": np.asarray(self._d, float),
                "uid": np.asarray(self._uid, np.uint64),
            }
        else:
            self._sync_from_device()
        return self._arrays

    # ------------------------------------------------------------------ device
    def _params(self):
        p = _abi.Params()
        p.max_iterations = self._max_iterations
        p.relaxation_model = self._model
        p.erp = self._erp
        p.abort_threshold = self._abort
        p.gravity[0], p.gravity[1], p.gravity[2] = self._gravity
        p.damping = self._damping
        p.relaxation_param = self._relax_param
        p.max_sub_iterations = self._max_sub
        p.sweep_mode = self._sweep_mode
        p.exact_trig = self._exact_trig
        p.sweep_schedule = self._schedule
        return p

    def _ensure_device(self):
        L = _abi.load()
        if self._ctx is None:
            h = ctypes.c_void_p()
            p = self._params()
            check(L.pehip_create(ctypes.byref(p), self.device, ctypes.byref(h)), None)
            self._ctx = h
            self._params_dirty = False
        if self._params_dirty:
            p = self._params()
            check(L.pehip_set_params(self._ctx, ctypes.byref(p)), self._ctx)
            self._params_dirty = False
        if self._bodies_dirty:
            a = self.host_arrays()
            n = len(a["shape"])
            if n == 0:
                raise ValueError("no bodies")
            nm = len(self.materials)
            csf = np.ascontiguousarray(self._csf, float)
            cdf = np.ascontiguousarray(self._cdf, float)
            check(L.pehip_set_materials(self._ctx, nm, ptr(csf), ptr(cdf)), self._ctx)
            c = {k: np.ascontiguousarray(v) for k, v in a.items()}
            check(L.pehip_upload_bodies(self._ctx, n, ptr(c["shape"], _abi._U8), ptr(c["sid"], _abi._U64),
                                        ptr(c["mat"], _abi._U32), ptr(c["fixed"], _abi._U8), ptr(c["param"]),
                                        ptr(c["pos"]), ptr(c["q"]), ptr(c["v"]), ptr(c["w"]), ptr(c["inv_mass"]),
                                        ptr(c["I"]), ptr(c["Iinv"]), ptr(c["normal"]), ptr(c["d"])), self._ctx)
            if self._m_owner:
                self._upload_members()
            self._bodies_dirty = False
            self._state_on_device = False

    def _upload_members(self):
        """Union members to the device (after their union bodies)."""
        k = len(self._m_owner)
        mo = np.ascontiguousarray(self._m_owner, np.uint32)
        ms = np.ascontiguousarray(self._m_sid, np.uint64)
        mm = np.ascontiguousarray(self._m_mat, np.uint32)
        mr = np.ascontiguousarray(self._m_rad, float)
        mrp = np.ascontiguousarray(np.asarray(self._m_rpos, float).reshape(k, 3))
        mg = np.ascontiguousarray(np.asarray(self._m_gpos, float).reshape(k, 3))
        check(_abi.load().pehip_upload_members(self._ctx, k, ptr(mo, _abi._U32), ptr(ms, _abi._U64),
                                               ptr(mm, _abi._U32), ptr(mr), ptr(mrp), ptr(mg)), self._ctx)

    def close(self):
        if self._ctx is not None:
            _abi.load().pehip_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ stepping
    def simulationStep(self, dt):
        """World::simulationStep (World.cpp:330-357)."""
        self.run(1, dt)

    def run(self, steps, dt):
        """World::run (World.cpp:374-378)."""
        if dt <= 0:
            raise ValueError("Invalid time step size")
        if self.size == 0:            # nothing to simulate: only the step counter advances
            self.timestep += int(steps)
            return
        self._ensure_device()
        check(_abi.load().pehip_step(self._ctx, float(dt), int(steps)), self._ctx)
        self._state_on_device = True
        self.timestep += int(steps)

    def detect(self):
        """Detection only on the current state; returns the contacts."""
        self._ensure_device()
        check(_abi.load().pehip_detect(self._ctx), self._ctx)
        return self.contacts()

    def setContactOrder(self, body1, body2):
        b1 = np.ascontiguousarray(body1, np.uint32).ravel()
        b2 = np.ascontiguousarray(body2, np.uint32).ravel()
        if len(b1) != len(b2):
            raise ValueError("contact order: %d first bodies but %d second bodies" % (len(b1), len(b2)))
        self._ensure_device()
        check(_abi.load().pehip_set_contact_order(self._ctx, len(b1), ptr(b1, _abi._U32), ptr(b2, _abi._U32)), self._ctx)

    def setForces(self, force, torque):
        f = _per_body(force, self.size, 3, "force")
        t = _per_body(torque, self.size, 3, "torque")
        self._ensure_device()
        check(_abi.load().pehip_set_forces(self._ctx, self.size, ptr(f), ptr(t)), self._ctx)

    # ------------------------------------------------------------------ state
    def getState(self):
        self._ensure_device()
        n = self.size
        st = {"pos": np.zeros((n, 3)), "q": np.zeros((n, 4)), "v": np.zeros((n, 3)), "w": np.zeros((n, 3)),
              "normal": np.zeros((n, 3)), "d": np.zeros(n)}
        check(_abi.load().pehip_download_state(self._ctx, n, ptr(st["pos"]), ptr(st["q"]), ptr(st["v"]),
                                               ptr(st["w"]), ptr(st["normal"]), ptr(st["d"])), self._ctx)
        if self._m_owner:   # union members' global positions
            st["mgpos"] = np.zeros((len(self._m_owner), 3))
            check(_abi.load().pehip_download_members(self._ctx, len(self._m_owner), ptr(st["mgpos"])), self._ctx)
        return st

    def setState(self, st):
        n = self.size
        c = {k: _per_body(st[k], n, w, k) for k, w in (("pos", 3), ("q", 4), ("v", 3), ("w", 3), ("normal", 3),
                                                       ("d", 0))}
        if "mgpos" in st and self._m_owner:
            _per_body(st["mgpos"], len(self._m_owner), 3, "mgpos")
        self._ensure_device()
        check(_abi.load().pehip_upload_state(self._ctx, n, ptr(c["pos"]), ptr(c["q"]), ptr(c["v"]), ptr(c["w"]),
                                             ptr(c["normal"]), ptr(c["d"])), self._ctx)
        if self._arrays is not None:
            for k in c:
                self._arrays[k] = c[k].copy()
        if "mgpos" in st and self._m_owner:   # union members are placed as given
            self._m_gpos = list(np.asarray(st["mgpos"], float).reshape(-1, 3))
            self._upload_members()
        self._state_on_device = False

    def contacts(self):
        L = _abi.load()
        n = ctypes.c_size_t()
        check(L.pehip_get_contacts(self._ctx, 0, ctypes.byref(n), None, None, None, None, None, None), self._ctx)
        nc = n.value
        c = {"b1": np.zeros(nc, np.uint32), "b2": np.zeros(nc, np.uint32), "pos": np.zeros((nc, 3)),
             "normal": np.zeros((nc, 3)), "dist": np.zeros(nc), "mu": np.zeros(nc)}
        check(L.pehip_get_contacts(self._ctx, nc, ctypes.byref(n), ptr(c["b1"], _abi._U32), ptr(c["b2"], _abi._U32),
                                   ptr(c["pos"]), ptr(c["normal"]), ptr(c["dist"]), ptr(c["mu"])), self._ctx)
        return c

    def sweepOrder(self):
        """(contact index per sweep slot, colour offsets) of the last step."""
        L = _abi.load()
        n = ctypes.c_size_t()
        check(L.pehip_get_sweep_order(self._ctx, 0, ctypes.byref(n), None, None, None), self._ctx)
        idx = np.zeros(n.value, np.uint32)
        ncol = ctypes.c_uint32()
        offs = np.zeros(130, np.uint32)
        check(L.pehip_get_sweep_order(self._ctx, n.value, ctypes.byref(n), ptr(idx, _abi._U32), ctypes.byref(ncol),
                                      ptr(offs, _abi._U32)), self._ctx)
        return idx, offs[:ncol.value + 1].copy()

    def impulses(self):
        L = _abi.load()
        n = ctypes.c_size_t()
        check(L.pehip_get_impulses(self._ctx, 0, ctypes.byref(n), None), self._ctx)
        p = np.zeros((n.value, 3))
        check(L.pehip_get_impulses(self._ctx, n.value, ctypes.byref(n), ptr(p)), self._ctx)
        return p

    def stats(self, kinetic=True):
        """Statistics of the last step; kinetic=False skips the kinetic-energy reduction (no
        kernel launch or synchronisation; kinetic_energy keeps its last computed value)."""
        s = _abi.Stats()
        if self._ctx is None:         # never stepped (e.g. a world without bodies): all zero
            return {f: getattr(s, f) for f, _ in _abi.Stats._fields_}
        fn = _abi.load().pehip_get_stats if kinetic else _abi.load().pehip_get_step_stats
        check(fn(self._ctx, ctypes.byref(s)), self._ctx)
        return {f: getattr(s, f) for f, _ in _abi.Stats._fields_}

    def setProfiling(self, on=True):
        self._ensure_device()
        check(_abi.load().pehip_set_profiling(self._ctx, 1 if on else 0), self._ctx)

    def evalTrig(self, fn, x, y=None, exact=True):
        """Device sin (0), cos (1), atan2(y, x) (2), acos (3), or sin (4) / cos (5) from one
        sincos, over an array."""
        self._ensure_device()
        xa = np.ascontiguousarray(x, float)
        ya = np.ascontiguousarray(y, float) if y is not None else None
        out = np.zeros_like(xa)
        check(_abi.load().pehip_eval_trig(self._ctx, int(fn), 1 if exact else 0, len(xa), ptr(xa), ptr(ya), ptr(out)),
              self._ctx)
        return out

    def stream(self):
        self._ensure_device()
        return _abi.load().pehip_stream(self._ctx)

    # CollisionSystem getters (HCSITS.h:715-737)
    def getMaximumPenetration(self):
        return self.stats()["max_penetration"]

    def getNumberOfContacts(self):
        return int(self.stats()["num_contacts"])
