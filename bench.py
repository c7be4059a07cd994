"""Benchmark of the HCSITS time step on one MI355X.

Workload (default): the polydisperse granular pile -- 200,000 spheres, radius U[0.25,0.5],
jittered lattice in a 60x60x200 m plane box, v0 ~ U[-0.5,0.5]^3, mu 0.5, 50 NBGS sweeps of
InelasticCoulombContactByDecoupling, dt 5e-4, FP64.  A "step" is one full time step
(detection, colouring, contact caching, 50 sweeps, integration) with all state resident in
HBM.  Prints one JSON line.
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


def cpu_baseline(sc, state, seconds, max_steps):
    """The oracle (single-threaded CPU restatement of pe's step) on a bounded sample: steps
    from the GPU state at the start of the timed region, until `seconds` have elapsed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ow = O.OracleWorld(gravity=sc.gravity, max_iterations=sc.max_iterations, model=sc.model)
    sc.to_oracle(ow)
    ow.set_state(state)
    ow.step(sc.dt, 1)          # first step also builds the hash grids from scratch; untimed
    n = 0
    t0 = time.perf_counter()
    while n < max_steps and time.perf_counter() - t0 < seconds:
        ow.step(sc.dt, 1)
        n += 1
    t = time.perf_counter() - t0
    return {"value": n / t, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": "%d steps of the same workload from the GPU state at the start of the timed region "
                      "(oracle/pe_oracle.cpp, g++ -O3, no FMA, 1 thread), %.1f s" % (n, t)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--spheres", type=int, default=200_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="no per-launch HIP events in the timed region")
    args = ap.parse_args()
    world_size = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world_size > 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(multi-GPU sharding is a design note, not implemented)" % max(args.gpus, world_size), file=sys.stderr)
        sys.exit(2)

    import torch
    from pe_amd import World, scenes

    sc = scenes.config2(args.spheres)
    w = World(device=0)
    sc.to_world(w)
    w.setMaxIterations(sc.max_iterations)
    w.setRelaxationModel(sc.model)

    if args.warmup:
        w.run(args.warmup, sc.dt)
    torch.cuda.synchronize()
    state0 = w.getState()
    if not args.no_profile:
        w.setProfiling(True)
    contacts = []
    colours = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        w.run(1, sc.dt)
        s = w.stats()
        contacts.append(s["num_contacts"])
        colours.append(s["num_colours"])
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    st = w.stats()

    steps_per_s = args.steps / t
    ci = float(np.sum(contacts)) * sc.max_iterations
    out = {
        "metric": METRIC, "value": steps_per_s, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "polydisperse granular pile, %d spheres r~U[0.25,0.5] in a 60x60x200 m plane box, "
                               "mu 0.5, 50 NBGS sweeps (InelasticCoulombContactByDecoupling), dt 5e-4, "
                               "steps %d..%d of the run" % (args.spheres, args.warmup, args.warmup + args.steps),
                   "bodies": int(w.size), "sweeps_per_step": sc.max_iterations, "dt": sc.dt,
                   "parallelism": "single GPU"},
        "contact_iterations_per_s": ci / t,
        "mean_contacts": float(np.mean(contacts)), "mean_colours": float(np.mean(colours)),
    }
    if not args.no_profile and st["prof_sweep_launches"]:
        avg_ms = st["prof_sweep_kernel_ms"] / st["prof_sweep_launches"]
        bytes_per_launch = st["prof_sweep_bytes"] / st["prof_sweep_launches"]
        achieved = st["prof_sweep_bytes"] / (st["prof_sweep_kernel_ms"] * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                           "kernel": "k_sweep<3> (one colour of one NBGS sweep)",
                           "avg_launch_ms": avg_ms, "bytes_per_launch": bytes_per_launch,
                           "launches": int(st["prof_sweep_launches"])}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, state0, args.cpu_seconds, args.steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
