"""GPU parity of the HIP time step against the CPU oracle (restatement of pe's HCSITS step).

Bars: the broadphase contact set (pairs, points, normals, distances, friction) is bit-exact;
a step in pe's sweep order (reference-order mode) matches to 1e-12; a coloured step equals
sequential Gauss-Seidel in colour order to 1e-12; converged coloured and reference-order
solves agree to 1e-8; trajectory observables stay within 1%.  Every comparison starts from
an identical state (granular dynamics amplify 1e-16 differences over a trajectory).
"""
import numpy as np
import pytest

import oracle as O
from pe_amd import World, scenes, SWEEP_REFERENCE_ORDER, SWEEP_COLOURED, PeHipError
from pe_amd._abi import SPHERE

pytestmark = pytest.mark.gpu


def rel_err(a, b, floor=1e-30):
    """Norm-wise relative difference max|a-b| / max|b|."""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), floor))


def contact_map(c):
    out = {}
    for i in range(len(c["b1"])):
        key = (int(c["b1"][i]), int(c["b2"][i]))
        out.setdefault(key, []).append((tuple(c["pos"][i]), tuple(c["normal"][i]), float(c["dist"][i]), float(c["mu"][i])))
    return out


def oracle_for(sc, iters=None, model=3, abort=0.0, cr_trig=False):
    ow = O.OracleWorld(gravity=sc.gravity, max_iterations=sc.max_iterations if iters is None else iters,
                       model=model, abort_threshold=abort, cr_trig=cr_trig)
    sc.to_oracle(ow)
    return ow


def gpu_for(sc, iters=None, model=3, mode=SWEEP_COLOURED, abort=0.0, exact_trig=False):
    w = World()
    sc.to_world(w)
    w.setMaxIterations(sc.max_iterations if iters is None else iters)
    w.setRelaxationModel(model)
    w.setSweepMode(mode)
    w.setAbortThreshold(abort)
    w.setExactTrig(exact_trig)
    return w


def dense_scene():
    # 216 spheres on a cubic lattice at spacing 0.9995 (overlaps ~5e-4) resting on the floor,
    # small random velocities: ~600 contacts, mostly static with some sliding
    return scenes.drop_scene(6, 6, 6, r=0.5, spacing=1.0 - 2e-6, jitter=2e-7, box=(7.5, 7.5, 20.0), mu=0.4,
                             seed=11, v_amp=0.05, max_iterations=30, z0=0.5 - 1e-6)


def poly_scene():
    # bidisperse packing: r = 0.5 on a cubic lattice plus r = 0.2073 spheres in the centres of
    # the horizontal lattice faces, each touching four big spheres.  Radius ratio 2.4 gives two
    # grid levels on the GPU and two HashGrids levels in pe.
    s = 1.0 - 2e-6
    base = scenes.drop_scene(5, 5, 4, r=0.5, spacing=s, jitter=2e-7, box=(7.0, 7.0, 20.0), mu=0.5, seed=5,
                             v_amp=0.05, max_iterations=30, z0=0.5 - 1e-6)
    o = np.array([0.5 + 0.5 * (s - 1.0) + 0.3, 0.5 + 0.5 * (s - 1.0) + 0.3, 0.5 - 1e-6])
    small = np.array([[o[0] + (i + 0.5) * s, o[1] + (j + 0.5) * s, o[2] + k * s]
                      for k in range(4) for j in range(4) for i in range(4)])
    rng = np.random.default_rng(6)
    small += rng.uniform(-2e-7, 2e-7, size=small.shape)
    base.sphere_pos = np.concatenate([base.sphere_pos, small])
    base.sphere_r = np.concatenate([base.sphere_r, np.full(len(small), s / np.sqrt(2) - 0.5 + 2e-6)])
    base.sphere_mat = np.zeros(len(base.sphere_pos), np.uint32)
    base.v0 = np.concatenate([base.v0, rng.uniform(-0.05, 0.05, size=(len(small), 3))])
    return base


def box_scene():
    # boxes resting on the floor (box-plane, 4 corners each) with spheres resting on their top
    # faces (sphere-box) and spheres on the floor between them (sphere-plane)
    rng = np.random.default_rng(9)
    nb = 6
    bpos = np.array([[1.0 + 1.6 * (k % 3), 1.2 + 2.2 * (k // 3), 0.4 - 1e-6] for k in range(nb)])
    ang = rng.uniform(-0.3, 0.3, size=nb)
    sph = []
    for k in range(nb):
        for dx in (-0.3005, 0.3005):
            sph.append([bpos[k][0] + dx, bpos[k][1] + 0.1, 0.8 - 1e-6 + 0.3 - 1e-6])
    for k in range(5):
        sph.append([0.5 + 1.1 * k, 2.3, 0.3 - 1e-6])
    sph = np.array(sph) + rng.uniform(-1e-7, 1e-7, size=(len(sph), 3))
    ns = len(sph)
    v0 = np.concatenate([np.zeros((6, 3)), rng.uniform(-0.3, 0.3, size=(ns, 3)) * [1, 1, 0], np.zeros((nb, 3))])
    return scenes.Scene("boxes", [{"name": "m", "density": 2500.0, "cor": 0.0, "csf": 0.3, "cdf": 0.3}],
                        scenes.box_planes(6.0, 6.0, 20.0), sph, np.full(ns, 0.3), np.zeros(ns, np.uint32),
                        box_pos=bpos, box_len=np.array([[1.0, 0.9, 0.8]] * nb),
                        box_q=np.array([[np.cos(a / 2), 0, 0, np.sin(a / 2)] for a in ang]),
                        box_mat=np.zeros(nb, np.uint32), v0=v0, dt=1e-3, max_iterations=30)


def _qaxis(axis, ang):
    ax = np.asarray(axis, float)
    return np.concatenate([[np.cos(ang / 2)], np.sin(ang / 2) * ax / np.linalg.norm(ax)])


def capsule_scene():
    # capsules lying on the floor (capsule-plane, both end spheres), parallel pairs side by side
    # (two capsule-capsule contacts), capsules crossed on top of others (one contact), spheres
    # resting on capsules (sphere-capsule), and capsules lying on boxes (box-capsule, from both
    # ends of the axis); small horizontal velocities
    rng = np.random.default_rng(31)
    e, r, L = 1e-6, 0.2, 0.8
    cpos, cq = [], []
    for k in range(4):                                          # floor row, parallel pairs
        for j in range(2):
            cpos.append([1.0 + 1.3 * k, 1.0 + j * (2 * r - e), r - e])
            cq.append([1.0, 0, 0, 0])
    for k in range(4):                                          # crossed on top of the pairs
        cpos.append([1.0 + 1.3 * k, 1.0 + 0.5 * (2 * r - e), 3 * r - 3 * e])
        cq.append(_qaxis((0, 0, 1), np.pi / 2 + rng.uniform(-0.2, 0.2)))
    bpos = np.array([[2.0, 4.0, 0.3 - e], [4.5, 4.0, 0.3 - e]])
    for b in bpos:                                              # lying on the boxes' top faces
        cpos.append([b[0], b[1], 0.6 - e + r - e])
        cq.append(_qaxis((0, 0, 1), rng.uniform(-0.5, 0.5)))
    cpos = np.array(cpos) + rng.uniform(-1e-7, 1e-7, size=(len(cpos), 3))
    nc = len(cpos)
    # spheres on top of the crossed capsules, over their centres
    sph = np.array([cpos[8 + k] + [0.0, 0.0, r + 0.25 - e] for k in range(4)])
    ns = len(sph)
    nb = len(bpos)
    v0 = np.concatenate([np.zeros((6, 3)), rng.uniform(-0.1, 0.1, size=(ns, 3)) * [1, 1, 0], np.zeros((nb, 3)),
                         rng.uniform(-0.2, 0.2, size=(nc, 3)) * [1, 1, 0]])
    return scenes.Scene("capsules", [{"name": "m", "density": 2000.0, "cor": 0.0, "csf": 0.3, "cdf": 0.3}],
                        scenes.box_planes(7.0, 6.0, 20.0), sph, np.full(ns, 0.25), np.zeros(ns, np.uint32),
                        box_pos=bpos, box_len=np.array([[1.4, 1.0, 0.6]] * nb), box_q=np.array([[1.0, 0, 0, 0]] * nb),
                        box_mat=np.zeros(nb, np.uint32), capsule_pos=cpos, capsule_r=np.full(nc, r),
                        capsule_len=np.full(nc, L), capsule_q=np.array(cq), capsule_mat=np.zeros(nc, np.uint32),
                        v0=v0, dt=1e-3, max_iterations=30)


def _granular(rng, R, mats):
    """Offsets, radii and materials of a granular particle (createGranularParticle's recipe,
    Union.cpp:2376-2402), and its spheres' offsets from its centre of mass."""
    k = int(rng.integers(2, 5))
    r = rng.uniform(0.6 * R, 0.8 * R, size=k)
    d = rng.uniform(-1.0, 1.0, size=(k, 3))
    d = d / np.linalg.norm(d, axis=1)[:, None] * (R - r)[:, None]
    mat = [int(mats[j % len(mats)]) for j in range(k)]
    m = r ** 3 * np.array([(2500.0, 1500.0)[x] for x in mat])
    com = (d * m[:, None]).sum(0) / m.sum()
    return [(d[j], float(r[j]), mat[j]) for j in range(k)], d - com, r


def _drop(xy, rel, rad, placed, e, top=0.0):
    """Height of a rigid cluster of spheres (offsets rel, radii rad) lowered at xy onto a flat
    support at height top and the spheres already placed (centre, radius), penetrating its
    support by e."""
    z = max(float(np.max(rad - rel[:, 2])) + top - e, 0.0)
    for o, ra in zip(rel, rad):
        for c, rb in placed:
            dxy = np.hypot(xy[0] + o[0] - c[0], xy[1] + o[1] - c[1])
            reach = ra + rb - e
            if dxy < reach:
                z = max(z, c[2] - o[2] + np.sqrt(reach * reach - dxy * dxy))
    return z


def union_scene():
    # granular particles (unions of 2-4 spheres) resting on the floor and on each other
    # (union-plane, union-union), spheres resting on them (sphere-union), particles on a box and
    # against a capsule (box-union, capsule-union); members of two materials
    rng = np.random.default_rng(47)
    e, R = 1e-6, 0.4
    placed = []
    bpos = np.array([[4.9, 1.5, 0.25]])
    cpos = np.array([[4.9, 3.2, 0.2 - e]])
    placed += [([4.3 + 0.005 * i, 3.2, 0.2 - e], 0.2) for i in range(241)]   # the capsule's axis
    unions = []

    def add(xy, mats, top=0.0):
        sph, rel, rad = _granular(rng, R, mats)
        z = _drop(xy, rel, rad, placed, e, top)
        unions.append(([xy[0], xy[1], z], sph))
        placed.extend((np.array([xy[0], xy[1], z]) + o, ra) for o, ra in zip(rel, rad))

    for iy in range(4):
        for ix in range(4):
            add((0.9 + 0.79 * ix, 0.9 + 0.79 * iy), (0, 1))
    for iy in range(3):
        for ix in range(3):
            add((1.3 + 0.79 * ix, 1.3 + 0.79 * iy), (1, 0))
    add((4.9, 1.5), (0,), top=0.5)                      # on the box's top face
    for k in range(2):
        add((4.5 + 0.8 * k, 3.75), (1,))
    spos = []
    for xy in ((1.7, 1.7), (2.5, 2.1), (2.1, 2.9)):
        z = _drop(xy, np.zeros((1, 3)), np.array([0.3]), placed, e)
        spos.append([xy[0], xy[1], z])
        placed.append((np.array(spos[-1]), 0.3))
    spos = np.array(spos)
    ns = len(spos)
    v0 = np.concatenate([np.zeros((6, 3)), rng.uniform(-0.05, 0.05, size=(ns, 3)) * [1, 1, 0], np.zeros((2, 3))])
    return scenes.Scene("unions", [{"name": "a", "density": 2500.0, "cor": 0.0, "csf": 0.2, "cdf": 0.2},
                                   {"name": "b", "density": 1500.0, "cor": 0.0, "csf": 0.1, "cdf": 0.1}],
                        scenes.box_planes(6.0, 6.0, 20.0), spos, np.full(ns, 0.3), np.ones(ns, np.uint32),
                        box_pos=bpos, box_len=np.array([[1.0, 1.0, 0.5]]), box_q=np.array([[1.0, 0, 0, 0]]),
                        box_mat=np.zeros(1, np.uint32), capsule_pos=cpos, capsule_r=np.full(1, 0.2),
                        capsule_len=np.full(1, 1.2), capsule_q=np.array([[1.0, 0, 0, 0]]),
                        capsule_mat=np.zeros(1, np.uint32), v0=v0, unions=unions, dt=1e-3, max_iterations=30)


def stack_scene():
    # box-box contacts of every kind collideBoxBox distinguishes (MaxContacts.h:867-2268):
    # face-face between boxes turned about z (incident-face clipping), an edge resting on a
    # face, two crossed edges (edge-edge), and two boxes side by side (parallel faces); plus
    # box-plane corners and spheres on the top box of a tower
    rng = np.random.default_rng(21)
    e, s45 = 1e-6, np.sqrt(0.5)
    pos, lens, qs = [], [], []
    for k in range(3):                                          # tower, turned about z
        pos.append([1.5, 1.5, 0.25 + 0.5 * k - (k + 1) * e])
        lens.append([1.0, 1.0, 0.5])
        qs.append(_qaxis((0, 0, 1), rng.uniform(-0.25, 0.25)))
    pos.append([4.5, 1.5, 0.3 - e]); lens.append([1.2, 1.2, 0.6]); qs.append(_qaxis((0, 0, 1), 0.1))
    h = 0.25 * s45 + 0.25 * s45                                 # edge resting on a face
    pos.append([4.5, 1.5, 0.6 - e + h - e]); lens.append([1.0, 0.5, 0.5]); qs.append(_qaxis((1, 0, 0), np.pi / 4))
    h = 0.3 * s45 + 0.3 * s45                                   # crossed edges
    pos.append([5.5, 5.0, h - e]); lens.append([0.6, 1.0, 0.6]); qs.append(_qaxis((0, 1, 0), np.pi / 4))
    pos.append([5.5, 5.0, 3 * h - 2 * e]); lens.append([1.0, 0.6, 0.6]); qs.append(_qaxis((1, 0, 0), np.pi / 4))
    pos.append([2.0, 5.0, 0.5 - e]); lens.append([1.0, 1.0, 1.0]); qs.append([1.0, 0, 0, 0])  # side by side
    pos.append([3.0 - e, 5.0, 0.5 - e]); lens.append([1.0, 1.0, 1.0]); qs.append([1.0, 0, 0, 0])
    nb = len(pos)
    sph = np.array([[1.3, 1.5, 1.5 - 3 * e + 0.2 - e], [1.7, 1.6, 1.5 - 3 * e + 0.2 - e]])
    sph += rng.uniform(-1e-7, 1e-7, size=sph.shape)
    ns = len(sph)
    v0 = np.concatenate([np.zeros((6, 3)), rng.uniform(-0.2, 0.2, size=(ns, 3)) * [1, 1, 0],
                         rng.uniform(-0.05, 0.05, size=(nb, 3)) * [1, 1, 0]])
    return scenes.Scene("stack", [{"name": "m", "density": 2000.0, "cor": 0.0, "csf": 0.4, "cdf": 0.4}],
                        scenes.box_planes(8.0, 8.0, 20.0), sph, np.full(ns, 0.2), np.zeros(ns, np.uint32),
                        box_pos=np.array(pos), box_len=np.array(lens), box_q=np.array(qs),
                        box_mat=np.zeros(nb, np.uint32), v0=v0, dt=1e-3, max_iterations=30)


def _box_R(q):
    w_, x_, y_, z_ = q
    return np.array([[1 - 2 * (y_ * y_ + z_ * z_), 2 * (x_ * y_ - w_ * z_), 2 * (x_ * z_ + w_ * y_)],
                     [2 * (x_ * y_ + w_ * z_), 1 - 2 * (x_ * x_ + z_ * z_), 2 * (y_ * z_ - w_ * x_)],
                     [2 * (x_ * z_ - w_ * y_), 2 * (y_ * z_ + w_ * x_), 1 - 2 * (x_ * x_ + y_ * y_)]])


def hopper_scene():
    # the mixed-shape hopper's geometry at small scale: two planes inclined at +-30 deg meeting
    # a 0.8 m gap above the floor (inclined normals: every body is tested against them).  On
    # each slope, slots of a box (edges U[0.3, 0.8], random orientation) or a sphere
    # (r ~ U[0.2, 0.4]) resting on the plane with a small sphere resting on top of it; three
    # spheres in the gap on the floor; mu = 0.6, sliding down the slopes
    rng = np.random.default_rng(31)
    L = 10.0
    cx = L / 2
    s30, c30 = 0.5, np.sqrt(0.75)
    nl, nr = np.array([s30, 0.0, c30]), np.array([-s30, 0.0, c30])
    planes = [((0, 0, 1), 0.0, 0), (tuple(nl), s30 * (cx - 0.4), 0), (tuple(nr), -s30 * (cx + 0.4), 0),
              ((0, 1, 0), 0.0, 0), ((0, -1, 0), -L, 0), ((1, 0, 0), 0.0, 0), ((-1, 0, 0), -L, 0)]
    sph, rad, bpos, blen, bq = [], [], [], [], []
    e = 1e-6
    for side, nrm, x0 in ((0, nl, cx - 0.4), (1, nr, cx + 0.4)):
        up_slope = np.array([-c30 if side == 0 else c30, 0.0, s30])   # along the slope, away from the gap
        for k in range(3):
            for j in range(4):
                foot = np.array([x0, 1.2 + 1.5 * j, 0.0]) + (0.9 + 1.5 * k) * up_slope
                if k > 0 and (k + j + side) % 2 == 0:
                    lens = rng.uniform(0.3, 0.8, 3)
                    q = scenes.shoemake_quaternions(rng, 1)[0]
                    h = 0.5 * np.sum(np.abs(nrm @ _box_R(q)) * lens)   # extent of the box along n
                    c = foot + (h - e) * nrm
                    bpos.append(c)
                    blen.append(lens)
                    bq.append(q)
                else:
                    h = rng.uniform(0.2, 0.4)
                    c = foot + (h - e) * nrm
                    sph.append(c)
                    rad.append(h)
                r2 = rng.uniform(0.15, 0.25)                           # resting on top, along n
                sph.append(c + (h + r2 - e) * nrm + rng.uniform(-1e-7, 1e-7, 3))
                rad.append(r2)
    for j in range(3):
        sph.append([cx + rng.uniform(-0.1, 0.1), 1.5 + 2.0 * j, 0.2 - e])
        rad.append(0.2)
    sph, rad = np.array(sph), np.array(rad)
    ns = len(sph)
    return scenes.Scene("hopper", [{"name": "m", "density": 2500.0, "cor": 0.0, "csf": 0.3, "cdf": 0.3}],
                        planes, sph, rad, np.zeros(ns, np.uint32), box_pos=np.array(bpos), box_len=np.array(blen),
                        box_q=np.array(bq), box_mat=np.zeros(len(bpos), np.uint32), dt=2e-4, max_iterations=30)


SCENES = {"dense": dense_scene, "poly": poly_scene, "boxes": box_scene, "stack": stack_scene,
          "hopper": hopper_scene, "capsules": capsule_scene, "unions": union_scene}


def oracle_state_after(sc, steps, iters=30, cr_trig=False):
    ow = oracle_for(sc, iters=iters, cr_trig=cr_trig)
    if steps:
        ow.step(sc.dt, steps)
    return ow, ow.get_state()


# --------------------------------------------------------------------------- broadphase
@pytest.mark.parametrize("name", sorted(SCENES))
@pytest.mark.parametrize("steps", [0, 7, 25])
def test_contact_set_bitexact(name, steps):
    sc = SCENES[name]()
    ow, st = oracle_state_after(sc, steps)
    w = gpu_for(sc)
    w.setState(st)
    g = contact_map(w.detect())
    ex = contact_map(ow.detect(exhaustive=True))
    hg = contact_map(ow.detect(exhaustive=False))
    assert sum(len(v) for v in ex.values()) > 10
    assert g == ex, "GPU contact set differs from pe's ExhaustiveSearch"
    assert g == hg, "GPU contact set differs from pe's HashGrids"


def test_contact_set_bitexact_c1_late_state():
    sc = scenes.config1()
    ow = oracle_for(sc, iters=100)
    ow.step(sc.dt, 470)
    st = ow.get_state()
    w = gpu_for(sc)
    w.setState(st)
    g = contact_map(w.detect())
    ex = contact_map(ow.detect(exhaustive=True))
    assert len(ex) > 50 and g == ex


def test_contact_set_bitexact_beyond_the_staged_contacts():
    # The count pass stages up to 8 contacts per emitting body; a body that emits more is written
    # by a second walk.  A sphere with its 12 kissing neighbours (all one size, the centre created
    # first, so it emits every pair) and, nearby, spheres with fewer contacts: both paths in one
    # detection, against pe's ExhaustiveSearch and HashGrids.
    c0 = np.array([5.0, 5.0, 3.0])
    dirs = [(1, 1, 0), (1, -1, 0), (-1, 1, 0), (-1, -1, 0), (1, 0, 1), (1, 0, -1), (-1, 0, 1), (-1, 0, -1),
            (0, 1, 1), (0, 1, -1), (0, -1, 1), (0, -1, -1)]
    pos = [c0] + [c0 + 0.99999 * np.array(d, float) / np.sqrt(2.0) for d in dirs]
    pos += [np.array([2.0 + 0.99999 * k, 2.0, 0.5 - 1e-6]) for k in range(4)]     # a row on the floor
    pos = np.array(pos)
    n = len(pos)
    sc = scenes.Scene("kissing", [{"name": "m", "density": 2500.0, "cor": 0.0, "csf": 0.3, "cdf": 0.3}],
                      scenes.box_planes(10.0, 10.0, 20.0), pos, np.full(n, 0.5), np.zeros(n, np.uint32),
                      dt=1e-3, max_iterations=20)
    ow = oracle_for(sc)
    w = gpu_for(sc)
    g = contact_map(w.detect())
    ex = contact_map(ow.detect(exhaustive=True))
    hg = contact_map(ow.detect(exhaustive=False))
    centre = int(np.flatnonzero(w.host_arrays()["shape"] == SPHERE)[0])
    assert sum(len(v) for (a, b), v in ex.items() if centre in (a, b)) == 12
    assert g == ex and g == hg


# --------------------------------------------------------------------------- one step, pe's order
@pytest.mark.parametrize("name", sorted(SCENES))
@pytest.mark.parametrize("model", [3, 1, 0, 2, 4])
def test_reference_order_step_matches_oracle(name, model):
    sc = SCENES[name]()
    ow, st0 = oracle_state_after(sc, 12)
    ow.L.orc_set_relaxation_model(ow.h, model)
    ow.step(sc.dt)                                  # pe's step from st0, HashGrids order
    c = ow.contacts()
    st1 = ow.get_state()
    w = gpu_for(sc, model=model, mode=SWEEP_REFERENCE_ORDER)
    w.setState(st0)
    w.setContactOrder(c["b1"], c["b2"])
    w.simulationStep(sc.dt)
    g1 = w.getState()
    assert len(c["b1"]) > 10
    errs = {k: rel_err(g1[k], st1[k]) for k in ("v", "w", "pos", "q")}
    perr = rel_err(w.impulses()[np.argsort(order_index(w.contacts(), c))], ow.impulses())
    # Model 3 searches with sin/cos/atan2/acos: against the oracle on the host libm, whose results
    # are not always correctly rounded, the union scene's velocities differ by ~1e-12; against the
    # correctly rounded oracle (the next test) every scene holds 1e-12.
    assert max(errs.values()) <= (5e-12 if (model == 3 and name == "unions") else 1e-12), errs
    # Model 3's friction of a box standing on four corners is statically indeterminate: the
    # split of the impulse among the corners moves by ~1e-8 with the host libm's rare non
    # correctly rounded trig results, while the velocities (the bar) do not.
    assert perr <= (1e-12 if model != 3 else 1e-7), perr


@pytest.mark.parametrize("model", [3, 5])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_reference_order_step_matches_correctly_rounded_oracle(name, model):
    # Against the oracle with correctly rounded trig (quad precision), the device's correctly
    # rounded trig leaves nothing to chance: velocities and every impulse agree to 1e-12.
    # Model 5 (generalized maximum dissipation) searches with trig as model 3 does.
    sc = SCENES[name]()
    ow, st0 = oracle_state_after(sc, 12, cr_trig=True)
    ow.L.orc_set_relaxation_model(ow.h, model)
    ow.step(sc.dt)
    c = ow.contacts()
    st1 = ow.get_state()
    w = gpu_for(sc, model=model, mode=SWEEP_REFERENCE_ORDER)
    w.setState(st0)
    w.setContactOrder(c["b1"], c["b2"])
    w.simulationStep(sc.dt)
    g1 = w.getState()
    assert len(c["b1"]) > 10
    for k in ("v", "w", "pos", "q"):
        assert rel_err(g1[k], st1[k]) <= 1e-12, k
    assert rel_err(w.impulses()[np.argsort(order_index(w.contacts(), c))], ow.impulses()) <= 1e-12


def pair_ranks(c):
    """Rank of each contact among the contacts of its body pair, in detection (narrowphase) order."""
    seen = {}
    r = np.zeros(len(c["b1"]), np.uint32)
    for i, (a, b) in enumerate(zip(c["b1"], c["b2"])):
        k = (int(a), int(b))
        r[i] = seen.get(k, 0)
        seen[k] = int(r[i]) + 1
    return r


def order_index(gc, oc):
    """Position of each GPU contact in the oracle's order (pairs matched in sequence)."""
    pos = {}
    for i, (a, b) in enumerate(zip(oc["b1"], oc["b2"])):
        pos.setdefault((int(a), int(b)), []).append(i)
    used = {}
    out = np.zeros(len(gc["b1"]), int)
    for i, (a, b) in enumerate(zip(gc["b1"], gc["b2"])):
        k = (int(a), int(b))
        j = used.get(k, 0)
        out[i] = pos[k][j]
        used[k] = j + 1
    return out


def test_reference_order_c1_window():
    # C1 (1k spheres, K = 100) stepped on the GPU in pe's order from pe's states, steps 440..444
    sc = scenes.config1()
    ow = oracle_for(sc, iters=100)
    ow.step(sc.dt, 440)
    w = gpu_for(sc, iters=100, mode=SWEEP_REFERENCE_ORDER)
    for _ in range(4):
        st0 = ow.get_state()
        ow.step(sc.dt)
        c = ow.contacts()
        w.setState(st0)
        w.setContactOrder(c["b1"], c["b2"])
        w.simulationStep(sc.dt)
        g, o = w.getState(), ow.get_state()
        assert len(c["b1"]) > 30
        for k in ("v", "w", "pos"):
            assert rel_err(g[k], o[k]) <= 1e-12, k


# --------------------------------------------------------------------------- coloured mode
@pytest.mark.parametrize("model", [3, 5])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_coloured_step_equals_gs_in_colour_order(name, model):
    # Contacts of one colour share no non-fixed body, so a coloured sweep is exactly a
    # sequential Gauss-Seidel sweep in colour-major order; replay that order on the oracle.
    # Exact-trig mode is compared with the oracle in its correctly rounded trig mode: the
    # host libm's rare non correctly rounded results would flip golden-section decisions.
    sc = SCENES[name]()
    ow, st0 = oracle_state_after(sc, 10, cr_trig=True)
    ow.L.orc_set_relaxation_model(ow.h, model)
    w = gpu_for(sc, iters=30, model=model, exact_trig=True)
    w.setState(st0)
    w.simulationStep(sc.dt)
    g1 = w.getState()
    idx, offs = w.sweepOrder()
    gc = w.contacts()
    assert len(offs) >= 3 and offs[-1] == len(idx) == len(gc["b1"])
    ow.set_sweep_order_ranked(gc["b1"][idx], gc["b2"][idx], pair_ranks(gc)[idx])
    ow.step(sc.dt)
    o1 = ow.get_state()
    for k in ("v", "w", "pos", "q"):
        assert rel_err(g1[k], o1[k]) <= 1e-12, k


@pytest.mark.parametrize("model", [2, 4])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_coloured_projection_step_equals_gs_in_colour_order(name, model):
    # The orthogonal-projection models (HCSITS.h:2631-2723) in colour order: no trig, so the
    # device and the oracle agree without the correctly rounded mode.
    sc = SCENES[name]()
    ow, st0 = oracle_state_after(sc, 10)
    ow.L.orc_set_relaxation_model(ow.h, model)
    w = gpu_for(sc, iters=30, model=model)
    w.setState(st0)
    w.simulationStep(sc.dt)
    g1 = w.getState()
    idx, offs = w.sweepOrder()
    gc = w.contacts()
    ow.set_sweep_order_ranked(gc["b1"][idx], gc["b2"][idx], pair_ranks(gc)[idx])
    ow.step(sc.dt)
    o1 = ow.get_state()
    assert len(gc["b1"]) > 10
    for k in ("v", "w", "pos", "q"):
        assert rel_err(g1[k], o1[k]) <= 1e-12, k


@pytest.mark.parametrize("model", [3, 5])
@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_dataflow_sweeps_equal_colour_launches(name, exact, model):
    # All sweeps in one cooperative launch, each contact waiting only for its own bodies' earlier
    # updates: the same per-body update order as the colour launches, so the same bits.
    sc = SCENES[name]()
    _, st0 = oracle_state_after(sc, 10)
    out = []
    for schedule in (1, 2):
        w = gpu_for(sc, iters=30, model=model, exact_trig=exact)
        w.setSweepSchedule(schedule)
        w.setState(st0)
        w.run(3, sc.dt)
        out.append((w.getState(), w.impulses()))
        w.close()
    for k in ("pos", "q", "v", "w"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("model", [3, 5, 1, 0, 2, 4])
@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_island_sweeps_equal_colour_launches(name, exact, model):
    # Each island of the contact graph swept on its own (small ones packed into wavefronts, large
    # ones in a workgroup, the largest by dataflow), colour by colour within the island: the colour
    # launches' per-body update order, so the same bits.
    sc = SCENES[name]()
    _, st0 = oracle_state_after(sc, 10)
    out = []
    for schedule in (1, 4):
        w = gpu_for(sc, iters=30, model=model, exact_trig=exact)
        w.setSweepSchedule(schedule)
        w.setState(st0)
        for _ in range(3):
            w.simulationStep(sc.dt)
            assert w.stats(kinetic=False)["schedule_used"] == schedule
        out.append((w.getState(), w.impulses(), w.stats(kinetic=False)["last_delta"]))
        w.close()
    for k in ("pos", "q", "v", "w"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def test_islands_of_every_kind():
    # one scene with islands of every unit kind: single contacts and small clusters (wavefront
    # units of several size classes), a 150-contact cluster (a workgroup unit) and a 400-contact
    # packing (dataflow), against the colour launches bit for bit
    rng = np.random.default_rng(5)
    pos, rad = [], []
    # separate clusters of touching spheres on the floor: sizes 1..40 spheres in lines
    x = 1.0
    for m in (1, 1, 2, 3, 5, 8, 12, 20, 33, 40):
        for k in range(m):
            pos.append([x + k * (1.0 - 2e-6), 2.0, 0.5 - 1e-6])
            rad.append(0.5)
        x += m + 1.5
    # an 8x8 touching layer (112 contacts + 64 with the floor: a workgroup unit)
    for i in range(8):
        for j in range(8):
            pos.append([2.0 + i * (1.0 - 2e-6), 6.0 + j * (1.0 - 2e-6), 0.5 - 1e-6])
            rad.append(0.5)
    # a 6x6x4 touching block (348 contacts + 36 with the floor: dataflow)
    for i in range(6):
        for j in range(6):
            for k in range(4):
                pos.append([20.0 + i * (1.0 - 2e-6), 8.0 + j * (1.0 - 2e-6), 0.5 - 1e-6 + k * (1.0 - 2e-6)])
                rad.append(0.5)
    pos = np.array(pos) + rng.uniform(-1e-7, 1e-7, size=(len(pos), 3))
    v0 = np.concatenate([np.zeros((6, 3)), rng.uniform(-0.05, 0.05, size=(len(pos), 3))])
    sc = scenes.Scene("islands", [{"name": "g", "density": 2500.0, "cor": 0.0, "csf": 0.2, "cdf": 0.2}],
                      scenes.box_planes(x + 2.0, 16.0, 20.0), pos, np.array(rad), np.zeros(len(pos), np.uint32),
                      v0=v0, dt=1e-3, max_iterations=30)
    out = []
    for schedule in (1, 4):
        w = gpu_for(sc, iters=30, model=3)
        w.setSweepSchedule(schedule)
        counts = []
        for _ in range(4):
            w.simulationStep(sc.dt)
            assert w.stats(kinetic=False)["schedule_used"] == schedule
            counts.append(w.stats(kinetic=False)["num_contacts"])
        out.append((w.getState(), w.impulses(), counts))
        w.close()
    assert out[0][2][0] == 800 and out[0][2] == out[1][2]   # the first step holds every kind
    for k in ("pos", "q", "v", "w"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("model,abort", [(3, 0.0), (1, 0.0), (0, 0.0), (1, 1e-3), (3, 1e-3), (2, 0.0), (4, 1e-3),
                                         (5, 0.0), (5, 1e-3)])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_single_block_sweeps_equal_colour_launches(name, model, abort):
    # All sweeps in one workgroup with a barrier between colours and the abort test on the
    # device: the colour launches' computation, and the same number of sweeps when aborting.
    sc = SCENES[name]()
    _, st0 = oracle_state_after(sc, 10)
    out = []
    for schedule in (1, 3):
        w = gpu_for(sc, iters=60, model=model, abort=abort)
        w.setSweepSchedule(schedule)
        w.setState(st0)
        its = []
        for _ in range(3):
            w.simulationStep(sc.dt)
            s = w.stats()
            assert s["schedule_used"] == schedule
            its.append(s["iterations"])
        out.append((w.getState(), w.impulses(), its))
        w.close()
    for k in ("pos", "q", "v", "w"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


@pytest.mark.parametrize("model", [3, 1, 0])
def test_single_block_abort_equals_colour_launches(model):
    # A resting packing converges (static contacts): both schedules stop after the same sweep.
    sc = scenes.drop_scene(3, 3, 3, r=0.5, spacing=1.0 - 2e-6, jitter=2e-7, box=(6.0, 6.0, 10.0), mu=0.4, seed=2,
                           v_amp=0.0, max_iterations=4000, z0=0.5 - 1e-6)
    ow = oracle_for(sc, iters=4000, abort=1e-10)
    ow.step(sc.dt, 3)
    st0 = ow.get_state()
    out = []
    for schedule in (1, 3):
        w = gpu_for(sc, iters=4000, model=model, abort=1e-10)
        w.setSweepSchedule(schedule)
        w.setState(st0)
        w.simulationStep(sc.dt)
        s = w.stats()
        assert s["schedule_used"] == schedule and s["num_contacts"] >= 20
        out.append((w.getState(), w.impulses(), s["iterations"], s["last_delta"]))
        w.close()
    assert out[0][2] == out[1][2] < 4000 and out[0][3] == out[1][3] < 1e-10
    for k in ("pos", "q", "v", "w"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("name", sorted(SCENES))
def test_colouring_is_conflict_free(name):
    sc = SCENES[name]()
    _, st0 = oracle_state_after(sc, 10)
    w = gpu_for(sc, iters=2)
    w.setState(st0)
    w.simulationStep(sc.dt)
    idx, offs = w.sweepOrder()
    gc = w.contacts()
    fixed = w.host_arrays()["fixed"]
    assert sorted(idx.tolist()) == list(range(len(gc["b1"])))
    for k in range(len(offs) - 1):
        seg = idx[offs[k]:offs[k + 1]]
        bodies = [int(b) for b in np.concatenate([gc["b1"][seg], gc["b2"][seg]]) if not fixed[b]]
        assert len(bodies) == len(set(bodies)), "colour %d shares a body" % k
    assert w.stats()["num_colours"] == len(offs) - 1


def test_coloured_converged_matches_reference_converged():
    # Both solvers run to a tight impulse-change threshold from the same state: the
    # velocities agree to 1e-8 even though the sweep orders differ.
    # A resting 3x3x3 packing: static contacts, so the fixed point is unique and pe's NBGS
    # converges (sliding configurations of the decoupled model do not converge at all).
    sc = scenes.drop_scene(3, 3, 3, r=0.5, spacing=1.0 - 2e-6, jitter=2e-7, box=(6.0, 6.0, 10.0), mu=0.4, seed=2,
                           v_amp=0.0, max_iterations=4000, z0=0.5 - 1e-6)
    ow = oracle_for(sc, iters=4000, abort=1e-10)
    ow.step(sc.dt, 3)
    st0 = ow.get_state()
    ow.step(sc.dt)
    o1 = ow.get_state()
    assert ow.stats()["num_contacts"] >= 20 and ow.stats()["iterations"] < 4000
    w = gpu_for(sc, iters=4000, abort=1e-10)
    w.setState(st0)
    w.simulationStep(sc.dt)
    g1 = w.getState()
    assert w.stats()["iterations"] < 4000
    for k in ("v", "w"):
        assert rel_err(g1[k], o1[k]) <= 1e-8, k


def test_converged_settled_c1_pile():
    # The 1k-sphere run continued to 4,000 steps settles into a pile (about 1,800 contacts).  Run
    # to pe's abort threshold (1e-7, HCSITS.h:424), the coloured solve and the oracle's solve in
    # pe's own order converge to the same impulses: velocities agree to 1e-8.  Only the
    # frictionless model converges there: on this state the oracle's decoupled and projected
    # friction models (1-4) stall at changes of 3e-3 to 0.4 after 20,000 sweeps (3: 4e-3 after
    # 100,000), so no converged solution exists to compare with.
    sc = scenes.config1()
    w = gpu_for(sc, iters=100, model=3)
    w.run(4000, sc.dt)
    st = w.getState()
    w.close()
    g = gpu_for(sc, iters=20000, model=0, abort=1e-7)
    g.setState(st)
    g.simulationStep(sc.dt)
    gs, gstats = g.getState(), g.stats()
    g.close()
    ow = oracle_for(sc, iters=20000, model=0, abort=1e-7)
    ow.set_state(st)
    ow.step(sc.dt)
    os_, ostats = ow.get_state(), ow.stats()
    assert gstats["num_contacts"] == ostats["num_contacts"] > 1500
    assert gstats["iterations"] < 20000 and ostats["iterations"] < 20000
    assert gstats["last_delta"] < 1e-7 and ostats["last_delta"] < 1e-7
    for k in ("v", "w"):
        assert rel_err(gs[k], os_[k]) <= 1e-8, (k, rel_err(gs[k], os_[k]))


def test_dataflow_at_the_automatic_limit():
    # about 60,000 contacts (a touching 30x30x22 packing), just under the automatic schedule's
    # dataflow limit of 65,536: the automatic choice is the dataflow launch, equal to the colour
    # launches bit for bit
    sc = scenes.drop_scene(30, 30, 22, r=0.5, spacing=1.0 - 2e-6, jitter=2e-7, box=(31.5, 31.5, 60.0), mu=0.4,
                           seed=21, v_amp=0.05, max_iterations=30, z0=0.5 - 1e-6)
    out = []
    for schedule in (0, 1):
        w = gpu_for(sc, iters=30, model=3)
        w.setSweepSchedule(schedule)
        w.simulationStep(sc.dt)
        s = w.stats(kinetic=False)
        out.append((w.getState(), w.impulses(), s))
        w.close()
    assert 50000 < out[0][2]["num_contacts"] <= 65536
    assert out[0][2]["schedule_used"] == 2 and out[1][2]["schedule_used"] == 1
    for k in ("pos", "q", "v", "w"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert np.array_equal(out[0][1], out[1][1])


def test_trajectory_observables_c1():
    # C1 over 500 steps: kinetic energy and maximum penetration of the coloured GPU run track
    # pe's sequential run within 1% (once contacts exist).
    sc = scenes.config1()
    ow = oracle_for(sc, iters=100)
    w = gpu_for(sc, iters=100)
    ke_o, ke_g, mp_o, mp_g = [], [], [], []
    m = 1.0 / sc.arrays()["inv_mass"][6:]
    for k in range(50):
        ow.step(sc.dt, 10)
        w.run(10, sc.dt)
        vo = ow.get_state()["v"][6:]
        ke_o.append(0.5 * np.sum(m * np.sum(vo * vo, axis=1)))
        s = w.stats()
        ke_g.append(s["kinetic_energy"])
        mp_o.append(ow.stats()["max_penetration"])
        mp_g.append(s["max_penetration"])
    ke_o, ke_g = np.array(ke_o), np.array(ke_g)
    rel = np.abs(ke_g - ke_o) / ke_o
    assert np.all(rel <= 0.01), "kinetic energy deviates: %s" % rel.max()
    assert abs(max(mp_g) - max(mp_o)) <= 0.01 * max(mp_o) + 1e-5, (max(mp_g), max(mp_o))


def test_coloured_run_is_deterministic():
    sc = dense_scene()
    runs = []
    for _ in range(2):
        w = gpu_for(sc, iters=20)
        w.run(15, sc.dt)
        runs.append(w.getState())
    for k in ("pos", "v", "w", "q"):
        assert np.array_equal(runs[0][k], runs[1][k])


def test_coloured_fast_trig_close_to_gs_in_colour_order():
    # With the device libm the golden-section decisions may differ from the host's in the
    # last bits, which moves velocities by ~1e-6 relative at most.
    sc = dense_scene()
    ow, st0 = oracle_state_after(sc, 10)
    w = gpu_for(sc, iters=30, exact_trig=False)
    w.setState(st0)
    w.simulationStep(sc.dt)
    g1 = w.getState()
    idx, _ = w.sweepOrder()
    gc = w.contacts()
    ow.set_sweep_order(gc["b1"][idx], gc["b2"][idx])
    ow.step(sc.dt)
    assert rel_err(g1["v"], ow.get_state()["v"]) <= 1e-5


def test_device_trig_matches_host_libm():
    # The correctly rounded device trig equals the host libm wherever the latter is correctly
    # rounded (glibc is not on ~0.1% of arguments); it never differs by more than one ulp.
    # (numpy's arctan2/arccos are not the C library's; the oracle calls libm, as math does)
    import math
    w = gpu_for(dense_scene())
    rng = np.random.default_rng(0)
    x = rng.uniform(-8.0, 8.0, 100_000)
    y = rng.uniform(-1.0, 1.0, 100_000)
    f = rng.uniform(0.0, 1.0, 100_000)
    host = lambda fn, *a: np.array([fn(*t) for t in zip(*a)])
    cases = [(0, x, None, host(math.sin, x)), (1, x, None, host(math.cos, x)),
             (2, x, y, host(math.atan2, y, x)), (3, f, None, host(math.acos, f))]
    for fn, a, b, host in cases:
        dev = w.evalTrig(fn, a, b, exact=True)
        mism = dev != host
        assert mism.mean() < 5e-3, (fn, mism.mean())
        ulp = np.abs(np.spacing(host))
        assert np.all(np.abs(dev - host) <= ulp), fn
    assert w.evalTrig(0, np.array([math.pi / 6]))[0] == math.sin(math.pi / 6)


def test_fast_sincos_within_two_ulp():
    # The production sin/cos (device libm) against the correctly rounded device trig.
    w = gpu_for(dense_scene())
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(-8.0, 8.0, 200_000), rng.uniform(-1e-3, 1e-3, 2000)])
    for fn in (0, 1):
        fast, cr = w.evalTrig(fn, x, exact=False), w.evalTrig(fn, x, exact=True)
        assert np.all(np.abs(fast - cr) <= 2 * np.abs(np.spacing(cr))), fn


@pytest.mark.parametrize("exact", [True, False])
def test_sincos_equals_sin_and_cos(exact):
    # The golden-section search takes sin and cos of one angle from a single sincos; it must
    # return exactly what separate sin and cos calls return.
    w = gpu_for(dense_scene())
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-8.0, 8.0, 200_000), rng.uniform(-1e-3, 1e-3, 1000), [0.0, -0.0, 1e-300]])
    assert np.array_equal(w.evalTrig(4, x, exact=exact), w.evalTrig(0, x, exact=exact))
    assert np.array_equal(w.evalTrig(5, x, exact=exact), w.evalTrig(1, x, exact=exact))
    assert np.array_equal(np.signbit(w.evalTrig(4, x, exact=exact)), np.signbit(w.evalTrig(0, x, exact=exact)))


# --------------------------------------------------------------------------- closed forms / errors
def test_free_fall_on_gpu():
    g, dt, n = 9.81, 1e-3, 40
    sc = scenes.drop_scene(1, 1, 1, r=0.5, box=(10, 10, 200), z0=100.0)
    w = gpu_for(sc)
    w.run(n, dt)
    st = w.getState()
    z0 = sc.sphere_pos[0][2]
    assert st["v"][6][2] == pytest.approx(-n * g * dt, rel=1e-13)
    assert st["pos"][6][2] == pytest.approx(z0 - g * dt * dt * n * (n + 1) / 2, rel=1e-13)


def test_resting_sphere_exact_on_gpu():
    sc = scenes.drop_scene(1, 1, 1, r=0.5, jitter=0.0, box=(10, 10, 20), z0=0.5)
    w = gpu_for(sc)
    w.simulationStep(1e-3)
    assert np.all(np.abs(w.getState()["v"][6]) < 1e-15)


def test_invalid_time_step_and_order():
    sc = dense_scene()
    w = gpu_for(sc, mode=SWEEP_REFERENCE_ORDER)
    with pytest.raises(ValueError):
        w.run(1, -1e-3)
    w.setContactOrder([0], [1])
    with pytest.raises(PeHipError):
        w.simulationStep(sc.dt)


def test_granular_pile_contacts_and_colour_step():
    # 1,500 granular particles (unions of 2-4 spheres, 12x12 a layer) after 1,200 oracle steps of
    # falling and piling: the GPU's contact set equals pe's HashGrids and ExhaustiveSearch bit for
    # bit, and one coloured step equals Gauss-Seidel in the same order on the oracle (correctly
    # rounded trig for that step) to 1e-12
    sc = scenes.config_granular(1500, per=12)
    ow = oracle_for(sc, iters=20)
    ow.step(sc.dt, 1200)
    ow.L.orc_set_trig_mode(ow.h, 1)
    st0 = ow.get_state()
    w = gpu_for(sc, iters=20, exact_trig=True)
    w.setState(st0)
    g = contact_map(w.detect())
    ex = contact_map(ow.detect(exhaustive=True))
    assert sum(len(v) for v in ex.values()) > 150
    assert g == ex and g == contact_map(ow.detect(exhaustive=False))
    w.simulationStep(sc.dt)
    g1 = w.getState()
    idx, offs = w.sweepOrder()
    gc = w.contacts()
    ow.set_state(st0)
    ow.set_sweep_order_ranked(gc["b1"][idx], gc["b2"][idx], pair_ranks(gc)[idx])
    ow.step(sc.dt)
    o1 = ow.get_state()
    for k in ("v", "w", "pos", "q", "mgpos"):
        assert rel_err(g1[k], o1[k]) <= 1e-12, k
