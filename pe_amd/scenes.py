"""Deterministic synthetic scenes for the time step (the benchmark workloads).

Each scene is a creation sequence in pe's terms: materials, then planes (as in
examples/lattice/Lattice.cpp:204-210, +z -z +y -y +x -x), then spheres and boxes.
Random numbers come from numpy's PCG64 seeded with 12345 + config id; lattice
sites are generated z-major, then y, then x.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from ._abi import BOX, CAPSULE, PLANE, SPHERE
from .world import granular_particle_spheres, mat3_inverse, normalize, quat_axis_angle, ACCURACY


@dataclass
class Scene:
    name: str
    materials: list                      # dicts: name, density, cor, csf, cdf
    planes: list                         # (normal, d, material)
    sphere_pos: np.ndarray = None        # (ns, 3)
    sphere_r: np.ndarray = None          # (ns,)
    sphere_mat: np.ndarray = None        # (ns,)
    box_pos: np.ndarray = None           # (nb, 3)
    box_len: np.ndarray = None           # (nb, 3)
    box_q: np.ndarray = None             # (nb, 4) orientation
    box_mat: np.ndarray = None
    capsule_pos: np.ndarray = None       # (nc, 3); capsules follow the boxes in creation order
    capsule_r: np.ndarray = None         # (nc,)
    capsule_len: np.ndarray = None       # (nc,) length of the cylinder part (body x-axis)
    capsule_q: np.ndarray = None         # (nc, 4) orientation
    capsule_mat: np.ndarray = None
    v0: np.ndarray = None                # (n_bodies, 3) initial linear velocities
    unions: list = None                  # created last, at rest: (gpos, [(offset, radius, material), ...])
    gravity: tuple = (0.0, 0.0, -9.81)
    dt: float = 1e-3
    steps: int = 1
    max_iterations: int = 100
    model: int = 3
    pair_friction: list = field(default_factory=list)

    @property
    def n_spheres(self):
        return 0 if self.sphere_pos is None else len(self.sphere_pos)

    @property
    def n_boxes(self):
        return 0 if self.box_pos is None else len(self.box_pos)

    @property
    def n_capsules(self):
        return 0 if self.capsule_pos is None else len(self.capsule_pos)

    @property
    def n_bodies(self):
        return len(self.planes) + self.n_spheres + self.n_boxes + self.n_capsules

    def arrays(self):
        """Body arrays (SoA) in creation order with pe's mass properties."""
        dens = np.array([m["density"] for m in self.materials])
        npl, ns, nb, nc = len(self.planes), self.n_spheres, self.n_boxes, self.n_capsules
        n = npl + ns + nb + nc
        a = {"shape": np.zeros(n, np.uint8), "mat": np.zeros(n, np.uint32), "fixed": np.zeros(n, np.uint8),
             "param": np.zeros((n, 3)), "pos": np.zeros((n, 3)), "q": np.zeros((n, 4)), "v": np.zeros((n, 3)),
             "w": np.zeros((n, 3)), "inv_mass": np.zeros(n), "I": np.zeros((n, 9)), "Iinv": np.zeros((n, 9)),
             "normal": np.zeros((n, 3)), "d": np.zeros(n)}
        a["q"][:, 0] = 1.0
        for k, (nrm, d, mat) in enumerate(self.planes):
            nv = normalize(nrm)
            if nv[0] * nv[0] + nv[1] * nv[1] < ACCURACY:
                q = quat_axis_angle((1.0, 0.0, 0.0), math.acos(nv[2]))
            else:
                q = quat_axis_angle((-nv[1], nv[0], 0.0), math.acos(nv[2]))
            a["shape"][k] = PLANE
            a["mat"][k] = mat
            a["fixed"][k] = 1
            a["pos"][k] = nv * float(d)
            a["q"][k] = q
            a["normal"][k] = nv
            a["d"][k] = float(d)
        if ns:
            s = slice(npl, npl + ns)
            r = np.asarray(self.sphere_r, float)
            rho = dens[np.asarray(self.sphere_mat)]
            m = 4.0 / 3.0 * math.pi * r * r * r * rho          # SphereBase.h:169-172
            I = np.zeros((ns, 9))
            I[:, 0] = I[:, 4] = I[:, 8] = 0.4 * m * r * r      # SphereBase.h:230-234
            a["shape"][s] = SPHERE
            a["mat"][s] = self.sphere_mat
            a["param"][s, 0] = r
            a["pos"][s] = self.sphere_pos
            a["inv_mass"][s] = 1.0 / m
            a["I"][s] = I
            a["Iinv"][s] = mat3_inverse(I)
        if nb:
            s = slice(npl + ns, npl + ns + nb)
            L = np.asarray(self.box_len, float)
            rho = dens[np.asarray(self.box_mat)]
            m = L[:, 0] * L[:, 1] * L[:, 2] * rho               # BoxBase.h:200-203
            I = np.zeros((nb, 9))
            I[:, 0] = m / 12.0 * (L[:, 1] * L[:, 1] + L[:, 2] * L[:, 2])   # BoxBase.cpp:152-158
            I[:, 4] = m / 12.0 * (L[:, 0] * L[:, 0] + L[:, 2] * L[:, 2])
            I[:, 8] = m / 12.0 * (L[:, 0] * L[:, 0] + L[:, 1] * L[:, 1])
            a["shape"][s] = BOX
            a["mat"][s] = self.box_mat
            a["param"][s] = L
            a["pos"][s] = self.box_pos
            a["q"][s] = self.box_q
            a["inv_mass"][s] = 1.0 / m
            a["I"][s] = I
            a["Iinv"][s] = mat3_inverse(I)
        if nc:
            s = slice(npl + ns + nb, n)
            r = np.asarray(self.capsule_r, float)
            L = np.asarray(self.capsule_len, float)
            rho = dens[np.asarray(self.capsule_mat)]
            m = r * r * math.pi * (4.0 / 3.0 * r + L) * rho     # CapsuleBase.cpp:60-92
            dn = m / (math.pi * r * r * (4.0 / 3.0 * r + L))     # CapsuleBase.h:196-199
            ms = 4.0 / 3.0 * math.pi * r * r * r * dn            # CapsuleBase.cpp:147-168
            mc = math.pi * r * r * L * dn
            I = np.zeros((nc, 9))
            I[:, 0] = r * r * (0.5 * mc + 0.4 * ms)
            I[:, 4] = I[:, 8] = mc * (0.25 * r * r + 1.0 / 12.0 * L * L) + ms * (0.4 * r * r + 0.375 * r * L + 0.25 * L * L)
            a["shape"][s] = CAPSULE
            a["mat"][s] = self.capsule_mat
            a["param"][s, 0] = r
            a["param"][s, 1] = L
            a["pos"][s] = self.capsule_pos
            a["q"][s] = self.capsule_q
            a["inv_mass"][s] = 1.0 / m
            a["I"][s] = I
            a["Iinv"][s] = mat3_inverse(I)
        if self.v0 is not None:
            a["v"][:] = self.v0
        return a

    def to_world(self, world):
        """Creates this scene in a pe_amd.World (bulk) and applies its solver settings."""
        for m in self.materials:
            world.createMaterial(m["name"], m["density"], m.get("cor", 0.0), m["csf"], m["cdf"])
        for (m1, m2, csf, cdf) in self.pair_friction:
            world.setPairFriction(m1, m2, csf, cdf)
        world.setGravity(*self.gravity)
        world.setMaxIterations(self.max_iterations)
        world.setRelaxationModel(self.model)
        world.add_bodies(self.arrays())
        for gpos, spheres in (self.unions or []):
            world.createUnion(0, gpos, spheres)
        return world

    def to_oracle(self, ow):
        """Creates this scene in the CPU oracle (one create call per body, in order)."""
        for m in self.materials:
            ow.create_material(m["density"], m.get("cor", 0.0), m["csf"], m["cdf"])
        for (m1, m2, csf, cdf) in self.pair_friction:
            ow.set_pair_friction(m1, m2, csf, cdf)
        for (nrm, d, mat) in self.planes:
            ow.create_plane(nrm, d, mat)
        for k in range(self.n_spheres):
            ow.create_sphere(self.sphere_pos[k], self.sphere_r[k], int(self.sphere_mat[k]))
        for k in range(self.n_boxes):
            ow.create_box(self.box_pos[k], self.box_len[k], int(self.box_mat[k]))
        for k in range(self.n_capsules):
            ow.create_capsule(self.capsule_pos[k], self.capsule_r[k], self.capsule_len[k], int(self.capsule_mat[k]))
        if self.v0 is not None or self.n_boxes or self.n_capsules:
            st = ow.get_state()
            if self.v0 is not None:
                st["v"][:] = self.v0
            b0 = len(self.planes) + self.n_spheres
            if self.n_boxes:
                st["q"][b0:b0 + self.n_boxes] = self.box_q
            if self.n_capsules:
                st["q"][b0 + self.n_boxes:] = self.capsule_q
            ow.set_state(st)
        for gpos, spheres in (self.unions or []):
            ow.create_union(gpos, [o for o, _, _ in spheres], [r for _, r, _ in spheres], [m for _, _, m in spheres])
        return ow


def box_planes(lx, ly, lz, mat=0):
    """Six inward planes bounding [0,lx]x[0,ly]x[0,lz] (Lattice.cpp:204-210 order)."""
    return [((0, 0, 1), 0.0, mat), ((0, 0, -1), -lz, mat), ((0, 1, 0), 0.0, mat), ((0, -1, 0), -ly, mat),
            ((1, 0, 0), 0.0, mat), ((-1, 0, 0), -lx, mat)]


def lattice(nx, ny, nz, spacing, origin, jitter, rng, count=None):
    """Jittered lattice sites, z-major then y then x; the first `count` sites."""
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = np.stack([origin[0] + spacing * i, origin[1] + spacing * j, origin[2] + spacing * k], axis=-1).reshape(-1, 3)
    if count is not None:
        base = base[:count]
    return base + rng.uniform(-jitter, jitter, size=base.shape)


def config1():
    """1,000 spheres r=0.5 on a jittered 10x10x10 lattice (spacing 1.2) in a 15x15x30 box;
    mu=0.4 (csf=cdf=0.2 per material), 100 sweeps, dt=1e-3, 500 steps."""
    rng = np.random.default_rng(12345 + 1)
    pos = lattice(10, 10, 10, 1.2, (1.5, 1.5, 1.0), 0.05, rng)
    n = len(pos)
    return Scene("C1 1k spheres", [{"name": "granular", "density": 2500.0, "cor": 0.0, "csf": 0.2, "cdf": 0.2}],
                 box_planes(15.0, 15.0, 30.0), pos, np.full(n, 0.5), np.zeros(n, np.uint32),
                 dt=1e-3, steps=500, max_iterations=100)


def config2(n_spheres=200_000):
    """Polydisperse pile: r ~ U[0.25,0.5], jittered lattice (spacing 1.05, jitter 0.02, 57x57 per
    layer) in a 60x60x200 box, v0 ~ U[-0.5,0.5]^3, mu=0.5, 50 sweeps, dt=5e-4, 2,000 steps."""
    rng = np.random.default_rng(12345 + 2)
    per = 57
    nz = (n_spheres + per * per - 1) // (per * per)
    pos = lattice(per, per, nz, 1.05, (0.525, 0.525, 0.525), 0.02, rng, count=n_spheres)
    r = rng.uniform(0.25, 0.5, size=n_spheres)
    v = rng.uniform(-0.5, 0.5, size=(n_spheres, 3))
    v0 = np.concatenate([np.zeros((6, 3)), v])
    return Scene("C2 %d spheres" % n_spheres,
                 [{"name": "granular", "density": 2500.0, "cor": 0.0, "csf": 0.25, "cdf": 0.25}],
                 box_planes(60.0, 60.0, 200.0), pos, r, np.zeros(n_spheres, np.uint32), v0=v0,
                 dt=5e-4, steps=2000, max_iterations=50)


def config_granular(n_particles=50_000, radius=0.5, per=57):
    """Pile of pe's granular particles (unions of 2-4 spheres, createGranularParticle) on a
    jittered lattice (spacing 1.05, jitter 0.02, per x per sites a layer) in a 60x60x200 box,
    released at rest; mu = 0.5, 50 sweeps, dt = 5e-4."""
    rng = np.random.default_rng(12345 + 6)
    nz = (n_particles + per * per - 1) // (per * per)
    pos = lattice(per, per, nz, 1.05, (0.525, 0.525, 0.525), 0.02, rng, count=n_particles)
    unions = [(pos[k], granular_particle_spheres(rng, radius, 0)) for k in range(n_particles)]
    return Scene("granular %d particles" % n_particles,
                 [{"name": "granular", "density": 2500.0, "cor": 0.0, "csf": 0.25, "cdf": 0.25}],
                 box_planes(60.0, 60.0, 200.0), unions=unions, dt=5e-4, steps=2000, max_iterations=50)


def drop_scene(nx, ny, nz, r=0.5, spacing=1.2, jitter=0.05, box=(10.0, 10.0, 40.0), mu=0.4, seed=7,
               v_amp=0.0, dt=1e-3, max_iterations=100, z0=None):
    """Small jittered-lattice drop into a plane box (test scenes)."""
    rng = np.random.default_rng(seed)
    z0 = r + 0.5 * (spacing - 2 * r) if z0 is None else z0
    pos = lattice(nx, ny, nz, spacing, (r + 0.5 * (spacing - 2 * r) + 0.3, r + 0.5 * (spacing - 2 * r) + 0.3, z0),
                  jitter, rng)
    n = len(pos)
    v0 = None
    if v_amp:
        v0 = np.concatenate([np.zeros((6, 3)), rng.uniform(-v_amp, v_amp, size=(n, 3))])
    return Scene("drop %dx%dx%d" % (nx, ny, nz),
                 [{"name": "granular", "density": 2500.0, "cor": 0.0, "csf": mu / 2, "cdf": mu / 2}],
                 box_planes(*box), pos, np.full(n, r), np.zeros(n, np.uint32), v0=v0, dt=dt,
                 max_iterations=max_iterations)


def shoemake_quaternions(rng, n):
    """Uniform random unit quaternions (w, x, y, z), Shoemake's method."""
    u1, u2, u3 = rng.uniform(0.0, 1.0, size=(3, n))
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    return np.stack([b * np.cos(2 * np.pi * u3), a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                     b * np.sin(2 * np.pi * u3)], axis=1)


def config3(n_spheres=950_000, n_boxes=50_000):
    """Mixed shapes filling a V-channel hopper: spheres r ~ U[0.2,0.4] and boxes with edges
    U[0.3,0.8]^3 in Shoemake-uniform orientations, on a jittered lattice (spacing 0.85, jitter
    0.02) that fills the hopper from the bottom and is poured into it: v0 = (0, 0, -10) +
    U[-1,1]^3 m/s.  The job lasts 0.2 s; a fill at rest or stirred only forms a dilute gas of
    brief contacts in that time (under 2,000 contacts at once), while the pour packs a layer
    several bodies deep against the inclined planes and the floor; mu = 0.6, 30 sweeps,
    dt = 2e-4, 1,000 steps.  The hopper: two planes inclined at +-30 deg whose lines meet 0.4 m
    either side of the channel axis at z = 0 (a 0.8 m outlet gap at the floor; pe planes are
    half-spaces, so the outlet is the gap rather than an opening), the floor, and four side walls
    of a 120 x 120 m footprint.  Boxes sit on the even sites of the lattice (every other site
    along each axis, spread evenly over the fill) with their six face neighbours left empty, so
    no two bodies overlap at the start: sphere-sphere face neighbours are >= 0.81 m apart
    (radii sum <= 0.8), a box's nearest spheres (edge neighbours) >= 1.14 m (extents sum
    <= 0.693 + 0.4) and two boxes >= 1.58 m (<= 1.386); sites keep 0.43 m (spheres) or 0.73 m
    (boxes) from every plane."""
    rng = np.random.default_rng(12345 + 3)
    L, sp, jit = 120.0, 0.85, 0.02
    c30, s30 = math.cos(math.radians(30.0)), math.sin(math.radians(30.0))
    cx = L / 2
    n_left = (s30, 0.0, c30)        # plane through (cx - 0.4, 0) rising to the left
    n_right = (-s30, 0.0, c30)
    planes = [((0, 0, 1), 0.0, 0),
              (n_left, s30 * (cx - 0.4), 0), (n_right, -s30 * (cx + 0.4), 0),
              ((0, 1, 0), 0.0, 0), ((0, -1, 0), -L, 0), ((1, 0, 0), 0.0, 0), ((-1, 0, 0), -L, 0)]
    nx = int(L / sp)
    nz = 1
    while True:                       # enough layers for the bodies after the boxes' clearances
        k, j, i = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(nx), indexing="ij")
        c = np.stack([sp / 2 + sp * i, sp / 2 + sp * j, sp / 2 + sp * k], axis=-1).reshape(-1, 3)
        dmin = np.full(len(c), np.inf)
        for (nrm, d, _) in planes:
            dmin = np.minimum(dmin, c @ np.asarray(nrm, float) - d)
        ii, jj, kk = i.reshape(-1), j.reshape(-1), k.reshape(-1)
        cand = np.flatnonzero((dmin >= 0.73) & (ii % 2 == 0) & (jj % 2 == 0) & (kk % 2 == 0))
        if len(cand) >= n_boxes and (dmin >= 0.43).sum() - 7 * n_boxes >= n_spheres:
            break
        nz += 8
    top = nz - 1
    for _ in range(4):                # spread the boxes over the layers the spheres fill
        sel = cand[kk[cand] <= top]
        if len(sel) < n_boxes:
            sel = cand
        is_box = np.zeros(len(c), bool)
        is_box[sel[np.linspace(0, len(sel) - 1, n_boxes).astype(np.int64)]] = True
        empty = np.zeros(len(c), bool)
        bi, bj, bk = ii[is_box], jj[is_box], kk[is_box]
        for di, dj, dk in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            ni, nj, nk = bi + di, bj + dj, bk + dk
            ok = (ni >= 0) & (ni < nx) & (nj >= 0) & (nj < nx) & (nk >= 0) & (nk < nz)
            empty[(nk[ok] * nx + nj[ok]) * nx + ni[ok]] = True
        sph_sites = np.flatnonzero((dmin >= 0.43) & ~is_box & ~empty)[:n_spheres]
        top = int(kk[sph_sites[-1]])
    box_sites = np.flatnonzero(is_box)
    sph = c[sph_sites] + rng.uniform(-jit, jit, size=(len(sph_sites), 3))
    box = c[box_sites] + rng.uniform(-jit, jit, size=(len(box_sites), 3))
    v = rng.uniform(-1.0, 1.0, size=(len(sph) + len(box), 3))
    v[:, 2] -= 10.0
    v0 = np.concatenate([np.zeros((len(planes), 3)), v])
    return Scene("C3 %d spheres + %d boxes" % (len(sph), len(box)),
                 [{"name": "granular", "density": 2500.0, "cor": 0.0, "csf": 0.3, "cdf": 0.3}],
                 planes, sph, rng.uniform(0.2, 0.4, size=len(sph)), np.zeros(len(sph), np.uint32),
                 box_pos=box, box_len=rng.uniform(0.3, 0.8, size=(len(box), 3)),
                 box_q=shoemake_quaternions(rng, len(box)), box_mat=np.zeros(len(box), np.uint32),
                 v0=v0, dt=2e-4, steps=1000, max_iterations=30)


def config4(n_scenes=4096):
    """Batched scenes: scene sizes round(exp(U[ln 50, ln 2000])), r = 0.5 spheres on a jittered
    8x8-per-layer lattice (spacing 1.1, jitter 0.05) in a 10x10x40 plane box each, per-scene
    friction mu ~ U[0.2, 0.8] (material csf = cdf = mu/2, so the pair friction is mu), 40 sweeps,
    dt = 1e-3, 300 steps.  Returns (scenes, mu)."""
    rng = np.random.default_rng(12345 + 4)
    sizes = np.rint(np.exp(rng.uniform(math.log(50.0), math.log(2000.0), size=n_scenes))).astype(int)
    mus = rng.uniform(0.2, 0.8, size=n_scenes)
    out = []
    for k in range(n_scenes):
        ns = int(sizes[k])
        pos = lattice(8, 8, (ns + 63) // 64, 1.1, (1.15, 1.15, 0.6), 0.05, rng, count=ns)
        out.append(Scene("C4 scene %d" % k,
                         [{"name": "granular", "density": 2500.0, "cor": 0.0, "csf": mus[k] / 2, "cdf": mus[k] / 2}],
                         box_planes(10.0, 10.0, 40.0), pos, np.full(ns, 0.5), np.zeros(ns, np.uint32),
                         dt=1e-3, steps=300, max_iterations=40))
    return out, mus


def config5(n=37):
    """Differentiable step: n^3 spheres r = 0.4 dropped as a jittered block (spacing 0.82, jitter
    0.005, bottom layer 0.1 m above the floor) onto a floor plane, mu = 0.5, dt = 5e-4, 400 steps."""
    rng = np.random.default_rng(12345 + 5)
    pos = lattice(n, n, n, 0.82, (0.41, 0.41, 0.5), 0.005, rng)
    ns = len(pos)
    return Scene("C5 %d spheres" % ns, [{"name": "granular", "density": 2500.0, "cor": 0.0, "csf": 0.25, "cdf": 0.25}],
                 [((0, 0, 1), 0.0, 0)], pos, np.full(ns, 0.4), np.zeros(ns, np.uint32),
                 dt=5e-4, steps=400, max_iterations=100)
