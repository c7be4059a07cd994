"""Benchmark of the HCSITS time step on one MI355X.

Workload (default): the polydisperse granular pile -- 200,000 spheres, radius U[0.25,0.5],
jittered lattice (spacing 1.05) in a 60x60x200 m plane box, v0 ~ U[-0.5,0.5]^3, mu 0.5, 50 NBGS
sweeps of InelasticCoulombContactByDecoupling, dt 5e-4, FP64, a 2,000-step job.  A "step" is one
full time step (detection, colouring, contact caching, 50 sweeps, integration) with all state
resident in HBM.  Prints one JSON line.

Timed window: the job's last --steps steps (default: the whole job).  The steps before the
window run untimed, the last --warmup of them as warmup; the window ends with the job, where the
pile holds the most contacts.  A whole-job window warms up on a separate copy of the scene.  The
CPU baseline times the oracle (single-threaded restatement of pe's step) from the GPU states at
the window's start and end (and middle, for the whole job).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time steps/s and contact-iterations/s at 200k spheres FP64; batched scenes/s"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the sweep kernel from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this
# bench (scaled by the counters' reading for a copy kernel of known size)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_sweep_dataflow_c2_traffic.json")


def progress(msg):
    """One line on stderr per finished phase (the JSON result alone goes to stdout)."""
    print("bench.py: " + msg, file=sys.stderr, flush=True)


def cpu_step_time(sc, state, seconds, min_steps=1):
    """Mean wall time of oracle steps from `state` (the first step, which also builds the hash
    grids from scratch, is not timed); runs until `seconds` have elapsed (at least min_steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ow = O.OracleWorld(gravity=sc.gravity, max_iterations=sc.max_iterations, model=sc.model)
    sc.to_oracle(ow)
    ow.set_state(state)
    ow.step(sc.dt, 1)
    n = 0
    t0 = time.perf_counter()
    while n < min_steps or time.perf_counter() - t0 < seconds:
        ow.step(sc.dt, 1)
        n += 1
    return (time.perf_counter() - t0) / n, n


def cpu_baseline(sc, states, steps, seconds, whole, first, job):
    """Oracle steps/s over the timed steps.  Whole job: per-step times at its start, middle and end
    (from the GPU states there) integrated with the trapezoid rule.  A window: the mean of the
    per-step times from the GPU states at its start and end."""
    per = seconds / len(states)
    ts, ns = [], []
    for st in states:
        t, n = cpu_step_time(sc, st, per)
        ts.append(t)
        ns.append(n)
        progress("cpu baseline: %d oracle steps, %.3f s/step" % (n, t))
    if whole:
        total = steps * (ts[0] + 2 * ts[1] + ts[2]) / 4.0
        where = ("from the GPU states at steps 0, %d and %d of the job (%.3f, %.3f, %.3f s/step), integrated over the "
                 "%d-step job (trapezoid rule)" % (steps // 2, steps, ts[0], ts[1], ts[2], steps))
    else:
        total = steps * (ts[0] + ts[1]) / 2.0
        where = ("from the GPU states at steps %d and %d, the start and end of the timed window (%.3f, %.3f s/step), "
                 "mean over the window's %d steps" % (first, job, ts[0], ts[1], steps))
    return {"value": steps / total, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": "oracle/pe_oracle.cpp (g++ -O3, no FMA, 1 thread) timed for %s steps %s"
                      % ("/".join(str(n) for n in ns), where)}


def scene_state(scs, st, k):
    """Slice of a whole-batch state that belongs to scene k (scenes stored back to back)."""
    off = np.concatenate([[0], np.cumsum([sc.n_bodies for sc in scs])])
    return {key: v[off[k]:off[k + 1]].copy() for key, v in st.items()}


def _oracle_scene_steps(job):
    """CPU worker: oracle time per step of the given batched scenes from their states (first
    step untimed), until the time budget is spent.  Returns (seconds, steps) per scene."""
    scs_idx, states, seconds = job
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from pe_amd import scenes
    scs, _ = scenes.config4(4096)
    out = []
    per = seconds / len(scs_idx)
    for k, st in zip(scs_idx, states):
        sc = scs[k]
        ow = O.OracleWorld(gravity=sc.gravity, max_iterations=40, model=3)
        sc.to_oracle(ow)
        ow.set_state(st)
        ow.step(sc.dt, 1)
        n, t0 = 0, time.perf_counter()
        while n < 2 or time.perf_counter() - t0 < per:
            ow.step(sc.dt, 1)
            n += 1
        out.append(((time.perf_counter() - t0) / n, n))
    return out


def batched_c4(torch, args):
    """Batched scenes (4,096 independent sphere scenes of 50-2,000 spheres, 40 sweeps, 300 steps,
    per-scene friction): the persistent batch kernel replayed from a hipGraph, against the
    single-world path stepping one scene at a time (sampled: 8 random scenes, 10 steps each from
    the batch's state at step 150, extrapolated to 300 steps x 4,096 scenes), and against the
    oracle on the host's cores (a seeded sample of scenes from the same state, 1 process and one
    process per available CPU up to 16)."""
    from pe_amd import World, scenes
    from pe_amd.batch import BatchWorld
    scs, mus = scenes.config4(4096)
    bw = BatchWorld(scs, scene_friction=(mus, mus))
    tun = bw.autotune(trial_steps=3)
    bw.setTuning(use_graph=1)
    nb_all = int(sum(sc.n_bodies for sc in scs))
    # 300 steps in chunks of 10: the last step of each chunk's contact count and sweep bytes
    # stand for the chunk's steps in the algorithmic-bytes sum
    chunk, tb, mid, sum_bytes, sum_contacts = 10, 0.0, None, 0.0, 0.0
    for k in range(0, 300, chunk):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bw.run(chunk)
        torch.cuda.synchronize()
        tb += time.perf_counter() - t0
        s = bw.stats()
        sum_bytes += chunk * (s["sweep_bytes"] + 776.0 * nb_all + 616.0 * s["num_contacts"])
        sum_contacts += chunk * s["num_contacts"]
        if k + chunk == 150:
            mid = bw.getState()
    progress("batched: 4096 scenes x 300 steps in %.2f s" % tb)
    st = bw.stats()
    bw.close()
    rng = np.random.default_rng(0)
    per = []
    for k in rng.choice(len(scs), 8, replace=False):
        sc = scs[int(k)]
        w = World()
        sc.to_world(w)
        w.setMaxIterations(40)
        w.setRelaxationModel(3)
        w.setState(scene_state(scs, mid, int(k)))
        w.run(2, sc.dt)
        torch.cuda.synchronize()
        a = time.perf_counter()
        w.run(10, sc.dt)
        torch.cuda.synchronize()
        per.append((time.perf_counter() - a) / 10)
        w.close()
    unb = 1.0 / (float(np.mean(per)) * 300)
    out = {"workload": "4096 independent scenes, log-uniform 50-2000 spheres r 0.5 in 10x10x40 m plane boxes, "
                       "per-scene mu U[0.2,0.8], 40 NBGS sweeps, dt 1e-3, 300 steps",
           "scenes_per_s": 4096 / tb, "scene_steps_per_s": 4096 * 300 / tb, "seconds": tb,
           "unbatched_scenes_per_s": unb, "batched_over_unbatched": (4096 / tb) / unb,
           "tuning": tun, "contacts_last_step": int(st["num_contacts"]),
           "mean_contacts": sum_contacts / 300,
           "roofline": {"bound": "hbm", "achieved": sum_bytes / tb / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": sum_bytes / tb / 1e9 / HBM_PEAK_GBS, "bytes_per_step": sum_bytes / 300,
                        "formula": "per scene and step K*(296*Nc + 200*(2*Nc - Nfixed)) + 776*Nb + 616*Nc, contact "
                                   "counts sampled every 10 steps, / wall time of the 300 steps"},
           "sample": "unbatched: single-world path, one scene at a time, 8 random scenes x 10 steps from the "
                     "batch state at step 150 (%.3f ms/step mean), extrapolated" % (1e3 * float(np.mean(per)))}
    if not args.no_cpu:
        out["cpu_baseline"] = batched_cpu(scs, mid, args.batched_cpu_seconds)
    return out


def batched_cpu(scs, mid, seconds):
    """Oracle scenes/s over the 4,096 scenes: 32 seeded scenes (the smallest and the largest among
    them) timed from the batch state at step 150; per-scene step times scaled by body count to the
    whole batch.  One process, then one process per available CPU (up to 16), each running the
    whole sample at the same time: the all-cores rate is the sum of the processes' rates under
    that load (a share of the sample per process would scale a few scenes' body-count ratio up to
    the batch, which made the estimate depend on which scenes a share got)."""
    import multiprocessing as mp
    rng = np.random.default_rng(4096)
    sizes = np.array([sc.n_bodies for sc in scs])
    idx = sorted(set(rng.choice(len(scs), 30, replace=False).tolist()) | {int(sizes.argmin()), int(sizes.argmax())})
    states = [scene_state(scs, mid, k) for k in idx]
    res = _oracle_scene_steps((idx, states, seconds))
    # per-step time per body of the sample, scaled to the batch's bodies: seconds per batch step
    t_sample = sum(t for t, _ in res)
    b_sample = sum(int(sizes[k]) for k in idx)
    t_batch_step = t_sample * float(sizes.sum()) / b_sample
    one = 1.0 / (300 * t_batch_step)
    model, ncpu = host_cpu()
    P = max(1, min(16, ncpu or 1))
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(P) as pool:
        outs = pool.map(_oracle_scene_steps, [(idx, states, seconds)] * P)
    wall = time.perf_counter() - t0
    # every process stepped the whole sample concurrently: each one's rate under load, summed
    allc = sum(1.0 / (300 * sum(t for t, _ in o) * float(sizes.sum()) / b_sample) for o in outs)
    progress("batched cpu baseline: 1 core %.4f scenes/s, %d processes %.4f scenes/s (%.1f s)" % (one, P, allc, wall))
    return {"value": one, "unit": "scenes/s", "cores": 1, "kind": "port", "all_cores_value": allc,
            "all_cores": P, "host_cpu": model, "host_cpus_available": ncpu,
            "sample": "oracle/pe_oracle.cpp (1 thread per process) on %d seeded scenes (%d-%d bodies) from the batch "
                      "state at step 150, about %.0f s each way, per-step time scaled by body count to the 4,096 "
                      "scenes x 300 steps; all cores: %d processes, each on the whole sample at once, rates summed"
                      % (len(idx), int(sizes[idx].min()), int(sizes[idx].max()), seconds, P)}


def step_signature(w):
    """Hash of the last step's active set: every contact's body pair, mode (separating, static
    inside the friction cone, sliding on it) and the sign of its distance (pe scales only negative
    distances by the error-reduction parameter, HCSITS.h:1795-1798)."""
    import hashlib
    c = w.contacts()
    p = w.impulses()
    pn = np.einsum("ij,ij->i", p, c["normal"])
    pt = np.linalg.norm(p - pn[:, None] * c["normal"], axis=1)
    mode = np.where(np.all(p == 0, axis=1), 0, np.where(pt < c["mu"] * pn * (1 - 1e-9), 1, 2)).astype(np.uint8)
    mode = mode | (np.asarray(c["dist"] < 0, np.uint8) << 2)
    key = (np.asarray(c["b1"], np.uint64) << np.uint64(32)) | np.asarray(c["b2"], np.uint64)
    i = np.argsort(key, kind="stable")
    return hashlib.sha1(key[i].tobytes() + mode[i].tobytes()).hexdigest(), len(key)


def adjoint_c5(args):
    """Differentiable step at full size: 37^3 spheres dropped onto a floor, 400 steps, loss
    |COM(T) - x*|^2; gradients w.r.t. every initial sphere velocity and mu from the reverse
    pass (checkpoints every 20 steps), checked against central differences (h = 1e-6) on
    seeded random coordinates whose perturbed runs keep every step's active set."""
    import torch
    from pe_amd import World, scenes
    from pe_amd.adjoint import DifferentiableRun
    sc = scenes.config5()
    model, mu, H = 1, 0.5, 1e-6
    iters, abort = args.c5_sweeps, args.c5_abort
    npl = len(sc.planes)
    n = sc.n_spheres
    m = 1.0 / sc.arrays()["inv_mass"][npl:]
    v0 = np.zeros((n, 3))

    def plain(vv, mm):
        w = World()
        sc.to_world(w)
        w.setMaxIterations(iters)
        w.setRelaxationModel(model)
        w.setAbortThreshold(abort)
        w.setPairFriction(0, 0, mm, mm)
        w.host_arrays()["v"][npl:] = vv
        sig, ncon = [], []
        for _ in range(sc.steps):
            w.run(1, sc.dt)
            h, k = step_signature(w)
            sig.append(h)
            ncon.append(k)
        pos = w.getState()["pos"][npl:]
        w.close()
        return pos, sig, ncon

    def loss(pos, target):
        com = (m[:, None] * pos).sum(0) / m.sum()
        d = com - target
        return float(d @ d), d

    t0 = time.perf_counter()
    base, base_sig, ncon = plain(v0, mu)
    t_plain = time.perf_counter() - t0
    progress("c5 plain forward: %.1f s, max contacts %d, steps with contacts %d" %
             (t_plain, max(ncon), sum(1 for k in ncon if k)))
    target = (m[:, None] * base).sum(0) / m.sum() + np.array([0.3, -0.2, 0.1])
    run = DifferentiableRun(sc, sc.steps, checkpoint_every=20, model=model, max_iterations=iters,
                            abort_threshold=abort, adjoint_sweeps=args.c5_adjoint_sweeps,
                            adjoint_tol=args.c5_adjoint_tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = run.forward(v0, mu)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    assert np.array_equal(st["pos"][npl:], base), "checkpointed forward differs from the plain run"
    L, d = loss(st["pos"][npl:], target)
    gpos = np.zeros((npl + n, 3))
    gpos[npl:] = 2.0 * d[None, :] * (m / m.sum())[:, None]
    t0 = time.perf_counter()
    _, gv, _, gmu = run.backward(gpos)
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    progress("c5 checkpointed forward %.1f s, reverse pass %.1f s" % (t_fwd, t_bwd))
    stats = dict(run.last_stats)
    run.close()
    # coordinates: seeded random ones over all 3n initial velocities, plus some of spheres that
    # touch the floor at the end (the random ones are mostly spheres still in free fall)
    rng = np.random.default_rng(64)
    coords = [("mu", None)] + [(int(f) // 3, int(f) % 3) for f in rng.choice(3 * n, args.fd_coords, replace=False)]
    touching = np.flatnonzero(base[:, 2] < 0.4 + 1e-6)
    if len(touching) and args.fd_contact_coords:
        coords += [(int(i), int(k)) for i, k in zip(rng.choice(touching, args.fd_contact_coords, replace=False),
                                                    rng.integers(0, 3, args.fd_contact_coords))]
    if args.fd_range:
        a, b = (int(x) for x in args.fd_range.split(":"))
        coords = coords[a:b]
    rows, skipped = [], 0
    M = m.sum()
    s_free = 2.0 * float(np.linalg.norm(d)) * sc.steps * sc.dt / n
    for c in coords:
        runs, same = [], True
        for sgn in (1, -1):
            if c[0] == "mu":
                pos, sig, _ = plain(v0, mu + sgn * H)
            else:
                vp = v0.copy()
                vp[c] += sgn * H
                pos, sig, _ = plain(vp, mu)
            same = same and sig == base_sig
            runs.append(pos)
        if not same:
            skipped += 1
            progress("c5 FD %s: active set changed, not comparable" % (c,))
            continue
        # L+ - L- = (d+ - d-) . (d+ + d-), with d+ - d- summed from per-sphere differences
        # (most spheres' final positions are identical in both runs, so no cancellation)
        dd = (m[:, None] * (runs[0] - runs[1])).sum(0) / M
        ds = loss(runs[0], target)[1] + loss(runs[1], target)[1]
        g_fd = float(dd @ ds) / (2 * H)
        g = gmu if c[0] == "mu" else float(gv[npl + c[0], c[1]])
        # relative to the larger of the two, floored (for derivatives that vanish: a sphere whose
        # push is absorbed by the floor, mu for a block landing flat) at the derivative a single
        # free-flying sphere has, 2 |d| T / n, or for mu at the loss itself
        scale = max(abs(g_fd), abs(g), L if c[0] == "mu" else s_free)
        rows.append({"coord": "mu" if c[0] == "mu" else "v0[%d,%d]" % c, "adjoint": g, "fd": g_fd,
                     "rel": abs(g - g_fd) / scale if scale > 0 else 0.0,
                     "touching": bool(c[0] != "mu" and c[0] in set(touching.tolist()))})
        progress("c5 FD %s: adjoint %.12e fd %.12e rel %.2e" % (rows[-1]["coord"], g, g_fd, rows[-1]["rel"]))
    return {"workload": "differentiable step: %d spheres r 0.4 dropped as a jittered 37^3 block onto a floor, mu 0.5, "
                        "relaxation model 1 (ApproximateInelasticCoulombContactByDecoupling), %s, adjoint sweeps to "
                        "%g, dt 5e-4, 400 steps, checkpoints every 20 steps; loss |COM(T) - x*|^2"
                        % (n, ("up to %d sweeps per step to |dp| < %g" % (iters, abort)) if abort > 0 else
                           ("%d sweeps per step" % iters), args.c5_adjoint_tol),
            "loss": L, "forward_s": t_fwd, "reverse_s": t_bwd, "plain_forward_s": t_plain,
            "max_contacts": int(max(ncon)), "adjoint_stats": stats, "grad_mu": gmu,
            "grad_v0_maxabs": float(np.abs(gv[npl:]).max()), "fd_h": H, "fd_checked": len(rows),
            "fd_not_comparable": skipped, "fd_scale_floor": s_free,
            "fd_max_rel": max((r["rel"] for r in rows), default=None), "fd": rows}


def save_state(path, st):
    np.savez(path, **st)


def load_state(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def host_cpu():
    """Model name of the host CPU and the number of CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count()
    return model, ncpu


def step_bytes(nb, nc, sweep_bytes):
    """Algorithmic HBM bytes of one time step: the sweeps' bytes (296 per contact + 200 per
    non-fixed contact end, per sweep) + 776 per body (predictor 256 read + 144 written,
    integration 152 + 152, broadphase 72) + 616 per contact (pair 8, narrowphase body reads 64,
    raw contact 56 + 56, setup body reads 160, contact record 272)."""
    return sweep_bytes + 776.0 * nb + 616.0 * nc


WORKLOAD_TEXT = {
    "c2": "polydisperse granular pile, %(n)d spheres r~U[0.25,0.5] in a 60x60x200 m plane box, v0~U[-0.5,0.5]^3, "
          "mu 0.5, 50 NBGS sweeps (InelasticCoulombContactByDecoupling), dt 5e-4, %(job)d-step job",
    "granular": "50000 granular particles (unions of 2-4 spheres, r 0.3-0.4 in a 0.5 bounding sphere) on a 57x57 "
                "jittered lattice in a 60x60x200 m plane box, released at rest, mu 0.5, 50 NBGS sweeps, dt 5e-4, "
                "%(job)d-step job",
    "c3": "mixed shapes, 950k spheres r~U[0.2,0.4] + 50k boxes edges U[0.3,0.8] (Shoemake orientations) filling a "
          "V-channel hopper (planes at +-30 deg, 0.8 m gap) on a 0.85 m jittered lattice, poured at v0 = (0,0,-10) + "
          "U[-1,1]^3 m/s, mu 0.6, 30 NBGS sweeps, dt 2e-4, %(job)d-step job",
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps: the last STEPS steps of the job (default: the whole job)")
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed steps just before the timed window (on a separate copy of the scene when the "
                         "window is the whole job)")
    ap.add_argument("--job-steps", type=int, default=None, help="length of the job (default: the workload's)")
    ap.add_argument("--spheres", type=int, default=200_000)
    ap.add_argument("--workload", choices=["c2", "c3", "c5", "granular"], default="c2",
                    help="c2: the polydisperse pile (default); c3: 950k spheres + 50k boxes in a V-channel hopper, "
                         "30 sweeps, dt 2e-4; granular: 50k unions of spheres; c5: the differentiable step")
    ap.add_argument("--cpu-seconds", type=float, default=24.0)
    ap.add_argument("--fd-coords", type=int, default=64, help="c5: random v0 coordinates checked by central differences")
    ap.add_argument("--fd-contact-coords", type=int, default=16,
                    help="c5: further coordinates of spheres that touch the floor at the end")
    ap.add_argument("--fd-range", default=None,
                    help="c5: START:STOP slice of the checked coordinates (mu, the random ones, the touching ones)")
    ap.add_argument("--c5-sweeps", type=int, default=20000, help="c5: sweep cap per step")
    ap.add_argument("--c5-abort", type=float, default=1e-15,
                    help="c5: sweep threshold on |dp|; differences of runs stopped early carry the truncation "
                         "error / 2h, so the solves run close to the floating-point fixed point")
    ap.add_argument("--c5-adjoint-sweeps", type=int, default=20000, help="c5: adjoint sweep cap per step")
    ap.add_argument("--c5-adjoint-tol", type=float, default=1e-20,
                    help="c5: adjoint sweep threshold (absolute change of the adjoint impulses, whose scale is ~1e-5)")
    ap.add_argument("--schedule", type=int, default=0, choices=[0, 1, 2, 3, 4],
                    help="sweep schedule: 0 auto, 1 one launch per colour, 2 dataflow, 3 one workgroup, 4 islands")
    ap.add_argument("--exact-trig", action="store_true",
                    help="correctly rounded sin/cos/atan2/acos in the solver (pehip_params.exact_trig)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="no per-launch HIP events in the timed region")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-scenes measurement")
    ap.add_argument("--batched-cpu-seconds", type=float, default=12.0,
                    help="oracle seconds per CPU worker for the batched scenes' CPU baseline")
    ap.add_argument("--save-state", default=None, help="save the state at the end of the timed window (.npz)")
    ap.add_argument("--save-window-start", default=None, help="save the state at the start of the timed window (.npz)")
    ap.add_argument("--dump-maps", default=None,
                    help="write the process's memory map (/proc/self/maps) to this file after the result line, to "
                         "place addresses of a stack trace printed at exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="before the run, 4 launches of a plain elementwise kernel that reads and writes 256 MiB "
                         "(scale reference for hardware byte counters)")
    args = ap.parse_args()
    world_size = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world_size > 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(multi-GPU sharding is a design note, not implemented)" % max(args.gpus, world_size), file=sys.stderr)
        sys.exit(2)

    import torch
    from pe_amd import World, scenes
    from pe_amd._abi import SCHEDULE_DATAFLOW, SCHEDULE_ISLANDS

    if args.workload == "c5":
        print(json.dumps({"metric": "adjoint of the time step: forward + reverse seconds, FD agreement",
                          "n_gpus": 1, "dtype": "f64", "data": "synthetic", **adjoint_c5(args)}))
        return

    sc = (scenes.config2(args.spheres) if args.workload == "c2" else
          scenes.config_granular() if args.workload == "granular" else scenes.config3())
    job = args.job_steps or sc.steps
    K = job if args.steps is None else args.steps
    W = args.warmup
    whole = K >= job
    if args.copy_calibration:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    def make_world():
        w = World(device=0)
        sc.to_world(w)
        w.setMaxIterations(sc.max_iterations)
        w.setRelaxationModel(sc.model)
        w.setSweepSchedule(args.schedule)
        w.setExactTrig(args.exact_trig)
        return w

    w = make_world()
    if whole:
        # the window is the whole job from the initial state: warm up on a separate copy
        first = 0
        if W:
            ww = make_world()
            ww.run(W, sc.dt)
            torch.cuda.synchronize()
            ww.close()
            progress("%d warmup steps done (separate copy of the scene)" % W)
    else:
        # the window is the job's last K steps: advance the job untimed to its start
        first = job - K
        pre = max(0, first - W)
        t0 = time.perf_counter()
        for k in range(0, pre, 250):
            w.run(min(250, pre - k), sc.dt)
            progress("advanced to step %d of %d (untimed)" % (min(k + 250, pre), job))
        if first - pre:
            w.run(first - pre, sc.dt)
        torch.cuda.synchronize()
        progress("steps 0..%d advanced untimed in %.1f s (the last %d as warmup)"
                 % (first, time.perf_counter() - t0, first - pre))
    state0 = w.getState()
    if args.save_window_start:
        save_state(args.save_window_start, state0)
    if not args.no_profile:
        w.setProfiling(True)
    contacts, colours, solve_ms, sweep_bytes, phases = [], [], [], [], []
    half = K // 2
    t = 0.0
    state_mid = None
    for seg in (half, K - half):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(seg):
            w.run(1, sc.dt)
            s = w.stats(kinetic=False)      # host-side record of the step: no launch, no sync
            contacts.append(s["num_contacts"])
            colours.append(s["num_colours"])
            solve_ms.append(s["t_solve_ms"])
            sweep_bytes.append(s["sweep_bytes"])
            phases.append([s["t_detection_ms"], s["t_colouring_ms"], s["t_setup_ms"], s["t_solve_ms"],
                           s["t_integrate_ms"]])
        torch.cuda.synchronize()
        t += time.perf_counter() - t0
        if state_mid is None:
            state_mid = w.getState()           # between the two timed halves
        progress("%d of %d timed steps done" % (len(contacts), K))
    st = w.stats()
    state_end = w.getState()
    nbodies = int(w.size)
    if args.save_state:
        save_state(args.save_state, state_end)
    w.close()                                  # every HIP object released before the next phase
    torch.cuda.synchronize()

    steps_per_s = K / t
    ci = float(np.sum(contacts)) * sc.max_iterations
    solve_s = float(np.sum(solve_ms)) * 1e-3
    window = ("whole job: steps 0..%d from the initial state" % K if whole else
              "timed window: steps %d..%d, the last %d steps of the %d-step job (steps 0..%d advanced untimed, "
              "the last %d of them as warmup)" % (first, job, K, job, first, min(W, first)))
    out = {
        "metric": METRIC, "value": steps_per_s, "unit": "steps/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": 1e3 * t / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": (WORKLOAD_TEXT[args.workload] % {"n": args.spheres, "job": job}) + "; " + window,
                   "bodies": nbodies, "sweeps_per_step": sc.max_iterations, "dt": sc.dt,
                   "window_first_step": first, "window_steps": K, "job_steps": job,
                   "trig": "correctly rounded" if args.exact_trig else "device libm",
                   "parallelism": "single GPU"},
        "contact_iterations_per_s": ci / solve_s if (solve_s > 0 and ci > 0) else None,
        "mean_contacts": float(np.mean(contacts)), "max_contacts": int(np.max(contacts)),
        "mean_colours": float(np.mean(colours)),
    }
    if not any(contacts):
        out["value_note"] = "no contacts in the timed window: the steps did no NBGS work"
    if not args.no_profile:
        ph = np.mean(np.array(phases), axis=0)
        out["phase_ms"] = {"detection": ph[0], "colouring": ph[1], "setup": ph[2], "solve": ph[3], "integration": ph[4],
                           "timing": "HIP events on the context stream, mean over the timed steps"}
    if not args.no_profile and st["prof_sweep_launches"]:
        avg_ms = st["prof_sweep_kernel_ms"] / st["prof_sweep_launches"]
        bytes_per_launch = st["prof_sweep_bytes"] / st["prof_sweep_launches"]
        achieved = st["prof_sweep_bytes"] / (st["prof_sweep_kernel_ms"] * 1e-3) / 1e9
        kernel = {SCHEDULE_DATAFLOW: "k_sweep_dataflow<3, TrigFast, 6>",
                  SCHEDULE_ISLANDS: "k_sweep_islands<3, TrigFast>"}.get(st["schedule_used"],
                                                                        "k_sweep<3> (one colour of one NBGS sweep)")
        traffic, traffic_src = None, None
        if os.path.exists(TRAFFIC_FILE) and args.workload == "c2" and args.spheres == 200_000:
            with open(TRAFFIC_FILE) as f:
                tf = json.load(f)
            if tf.get("kernel") == kernel and tf.get("window") == [first, job]:
                traffic = tf.get("hbm_bytes_per_launch")
                traffic_src = tf.get("source")
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                           "traffic_source": traffic_src or "not measured for this window",
                           "kernel": kernel + (" (all %d NBGS sweeps of one step)" % sc.max_iterations
                                               if st["schedule_used"] in (SCHEDULE_DATAFLOW, SCHEDULE_ISLANDS)
                                               else ""),
                           "avg_launch_ms": avg_ms, "bytes_per_launch": bytes_per_launch,
                           "launches": int(st["prof_sweep_launches"]),
                           "timing": "HIP events around every sweep launch on the context stream"}
    total_bytes = sum(step_bytes(nbodies, c, b) for c, b in zip(contacts, sweep_bytes))
    out["step_roofline"] = {"bound": "hbm", "achieved": total_bytes / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": total_bytes / t / 1e9 / HBM_PEAK_GBS, "bytes_per_step": total_bytes / K,
                            "formula": "sum over the timed steps of K*(296*Nc + 200*(2*Nc - Nfixed)) + 776*Nb + 616*Nc "
                                       "bytes / wall time of the window"}
    if not args.no_batched and args.workload == "c2" and args.spheres == 200_000:
        out["batched"] = batched_c4(torch, args)
    if not args.no_cpu:
        model, ncpu = host_cpu()
        states = [state0, state_mid, state_end] if whole else [state0, state_end]
        out["cpu_baseline"] = cpu_baseline(sc, states, K, args.cpu_seconds, whole, first, job)
        out["cpu_baseline"]["host_cpu"] = model
        out["cpu_baseline"]["host_cpus_available"] = ncpu
    print(json.dumps(out))
    sys.stdout.flush()
    if args.dump_maps:
        with open("/proc/self/maps") as f, open(args.dump_maps, "w") as g:
            g.write(f.read())


if __name__ == "__main__":
    main()
