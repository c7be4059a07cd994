"""Batched scenes: many independent pe worlds advanced together on one GPU.

pe simulates one world per process (the World and CollisionSystem singletons,
src/core/World.cpp:65, HardContactSemiImplicitTimesteppingSolvers.h:118); a
``BatchWorld`` holds many such worlds ("scenes") and advances all of them with one
persistent kernel (``pehip_batch_*`` in include/pe_hip.h).  Every scene follows the
semantics of a ``World`` in coloured mode.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import check_batch, ptr
from .world import _per_body


class BatchWorld:
    def __init__(self, scenes, device=0, contacts_per_body=8.0, scene_friction=None, max_iterations=None,
                 relaxation_model=None, abort_threshold=0.0, exact_trig=False):
        """scenes: list of pe_amd.scenes.Scene.  All scenes share the solver settings of the first
        one (gravity, sweeps, relaxation model).  The scenes' materials go into one shared pair
        table (identical materials shared).  scene_friction: optional (static_mu, dynamic_mu)
        arrays, one value per scene, replacing the pair tables' friction."""
        L = _abi.load()
        self.scenes = scenes
        first = scenes[0]
        p = _abi.Params()
        L.pehip_default_params(ctypes.byref(p))
        p.max_iterations = first.max_iterations if max_iterations is None else max_iterations
        p.relaxation_model = first.model if relaxation_model is None else relaxation_model
        p.abort_threshold = abort_threshold
        p.gravity[0], p.gravity[1], p.gravity[2] = first.gravity
        p.exact_trig = 1 if exact_trig else 0
        self._params = p
        h = ctypes.c_void_p()
        check_batch(L.pehip_batch_create(ctypes.byref(p), device, ctypes.byref(h)), None)
        self._b = h
        self.dt = first.dt
        # shared material table: identical materials are shared between scenes; with per-scene
        # friction only the densities matter
        mats, index, maps = [], {}, []
        dedup = not any(sc.pair_friction for sc in scenes)
        for sc in scenes:
            local = []
            for m in sc.materials:
                key = (m["density"],) if scene_friction is not None else tuple(sorted(m.items()))
                if not dedup or key not in index:
                    index[key] = len(mats)
                    mats.append(m)
                local.append(index[key])
            maps.append(np.array(local, np.uint32))
        nm = len(mats)
        if nm > 4096:
            raise ValueError("more than 4096 distinct materials; use scene_friction for per-scene friction")
        csf = np.zeros((nm, nm))
        cdf = np.zeros((nm, nm))
        if scene_friction is None:
            for i in range(nm):
                for j in range(nm):
                    csf[i, j] = mats[i]["csf"] + mats[j]["csf"]   # Materials.cpp:237-243
                    cdf[i, j] = mats[i]["cdf"] + mats[j]["cdf"]
            for sc, mp in zip(scenes, maps):
                for (m1, m2, s_, d_) in sc.pair_friction:
                    csf[mp[m1], mp[m2]] = csf[mp[m2], mp[m1]] = s_
                    cdf[mp[m1], mp[m2]] = cdf[mp[m2], mp[m1]] = d_
        check_batch(L.pehip_batch_set_materials(self._b, nm, ptr(csf), ptr(cdf)), self._b)
        arrs = []
        for sc, mp in zip(scenes, maps):
            a = sc.arrays()
            a["mat"] = mp[a["mat"]]
            arrs.append(a)
        self.offsets = np.zeros(len(scenes) + 1, np.uint32)
        self.offsets[1:] = np.cumsum([len(a["shape"]) for a in arrs])
        cat = {k: np.ascontiguousarray(np.concatenate([a[k] for a in arrs])) for k in arrs[0]}
        cat["sid"] = np.arange(len(cat["shape"]), dtype=np.uint64)
        self.n = int(self.offsets[-1])
        check_batch(L.pehip_batch_upload(self._b, len(scenes), ptr(self.offsets, _abi._U32),
                                         ptr(cat["shape"], _abi._U8), ptr(cat["sid"], _abi._U64),
                                         ptr(cat["mat"], _abi._U32), ptr(cat["fixed"], _abi._U8), ptr(cat["param"]),
                                         ptr(cat["pos"]), ptr(cat["q"]), ptr(cat["v"]), ptr(cat["w"]),
                                         ptr(cat["inv_mass"]), ptr(cat["I"]), ptr(cat["Iinv"]), ptr(cat["normal"]),
                                         ptr(cat["d"]), float(contacts_per_body)), self._b)
        if scene_friction is not None:
            self.setSceneFriction(*scene_friction)

    @property
    def nscenes(self):
        return len(self.scenes)

    def setSceneFriction(self, static_mu, dynamic_mu):
        s = np.ascontiguousarray(static_mu, float)
        d = np.ascontiguousarray(dynamic_mu, float)
        check_batch(_abi.load().pehip_batch_set_scene_friction(self._b, self.nscenes, ptr(s), ptr(d)), self._b)

    def run(self, steps, dt=None):
        dt = self.dt if dt is None else dt
        check_batch(_abi.load().pehip_batch_step(self._b, float(dt), int(steps)), self._b)

    def tuning(self):
        t = _abi.BatchTuning()
        check_batch(_abi.load().pehip_batch_get_tuning(self._b, ctypes.byref(t)), self._b)
        return {k: getattr(t, k) for k, _ in t._fields_}

    def setTuning(self, **kw):
        cur = self.tuning()
        cur.update(kw)
        t = _abi.BatchTuning(**cur)
        check_batch(_abi.load().pehip_batch_set_tuning(self._b, ctypes.byref(t)), self._b)

    def autotune(self, trial_steps=5, dt=None):
        dt = self.dt if dt is None else dt
        t = _abi.BatchTuning()
        check_batch(_abi.load().pehip_batch_autotune(self._b, float(dt), int(trial_steps), ctypes.byref(t)), self._b)
        return {k: getattr(t, k) for k, _ in t._fields_}

    def getState(self):
        n = self.n
        st = {"pos": np.zeros((n, 3)), "q": np.zeros((n, 4)), "v": np.zeros((n, 3)), "w": np.zeros((n, 3)),
              "normal": np.zeros((n, 3)), "d": np.zeros(n)}
        check_batch(_abi.load().pehip_batch_download_state(self._b, n, ptr(st["pos"]), ptr(st["q"]), ptr(st["v"]),
                                                           ptr(st["w"]), ptr(st["normal"]), ptr(st["d"])), self._b)
        return st

    def setState(self, st):
        c = {k: _per_body(st[k], self.n, w, k) for k, w in (("pos", 3), ("q", 4), ("v", 3), ("w", 3), ("normal", 3),
                                                            ("d", 0))}
        check_batch(_abi.load().pehip_batch_upload_state(self._b, self.n, ptr(c["pos"]), ptr(c["q"]), ptr(c["v"]),
                                                         ptr(c["w"]), ptr(c["normal"]), ptr(c["d"])), self._b)

    def sceneState(self, st, k):
        """Slice of a full-batch state belonging to scene k."""
        lo, hi = int(self.offsets[k]), int(self.offsets[k + 1])
        return {key: v[lo:hi].copy() for key, v in st.items()}

    def sceneStats(self):
        m = self.nscenes
        out = {"contacts": np.zeros(m, np.uint32), "colours": np.zeros(m, np.uint32),
               "iterations": np.zeros(m, np.uint32), "max_penetration": np.zeros(m), "steps": np.zeros(m, np.uint32)}
        check_batch(_abi.load().pehip_batch_scene_stats(self._b, m, ptr(out["contacts"], _abi._U32),
                                                        ptr(out["colours"], _abi._U32),
                                                        ptr(out["iterations"], _abi._U32), ptr(out["max_penetration"]),
                                                        ptr(out["steps"], _abi._U32)), self._b)
        return out

    def contacts(self, k):
        """Contacts of scene k from its last step, body indices local to the scene."""
        L = _abi.load()
        n = ctypes.c_size_t()
        check_batch(L.pehip_batch_get_contacts(self._b, k, 0, ctypes.byref(n), None, None, None, None, None, None),
                    self._b)
        m = n.value
        c = {"b1": np.zeros(m, np.uint32), "b2": np.zeros(m, np.uint32), "pos": np.zeros((m, 3)),
             "normal": np.zeros((m, 3)), "dist": np.zeros(m), "mu": np.zeros(m)}
        check_batch(L.pehip_batch_get_contacts(self._b, k, m, ctypes.byref(n), ptr(c["b1"], _abi._U32),
                                               ptr(c["b2"], _abi._U32), ptr(c["pos"]), ptr(c["normal"]),
                                               ptr(c["dist"]), ptr(c["mu"])), self._b)
        c["b1"] -= self.offsets[k]
        c["b2"] -= self.offsets[k]
        return c

    def sweepOrder(self, k):
        L = _abi.load()
        n = ctypes.c_size_t()
        ncol = ctypes.c_uint32()
        check_batch(L.pehip_batch_get_sweep_order(self._b, k, 0, ctypes.byref(n), None, ctypes.byref(ncol), None),
                    self._b)
        idx = np.zeros(n.value, np.uint32)
        offs = np.zeros(ncol.value + 1, np.uint32)
        check_batch(L.pehip_batch_get_sweep_order(self._b, k, n.value, ctypes.byref(n), ptr(idx, _abi._U32),
                                                  ctypes.byref(ncol), ptr(offs, _abi._U32)), self._b)
        return idx, offs

    def stats(self):
        s = _abi.BatchStats()
        check_batch(_abi.load().pehip_batch_get_stats(self._b, ctypes.byref(s)), self._b)
        return {k: getattr(s, k) for k, _ in s._fields_}

    def stream(self):
        return _abi.load().pehip_batch_stream(self._b)

    def close(self):
        if getattr(self, "_b", None) is not None:
            _abi.load().pehip_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
