#!/usr/bin/env python3
"""bench.py -- body-steps/s of the MI355X step at 1M bodies.

Workload: the plain-method point of the headline metric, a periodic elastic
sphere gas of 1,048,576 spheres (r = 0.5, m = 1) at the 4096-sphere/L=40
density (L = 253.98), uniform positions rejection-sampled to no overlap
(seed 0), velocities N(0,1) with zero net momentum, e = 1, mu = 0, no gravity,
dt = 1/60, 8 solver iterations, FP64.  Inputs are synthetic and resident in
HBM before the timed region; a step is one full `World::step()`
(integrate, broadphase, narrow phase, warm-started sequential impulses,
positional correction, periodic wrap).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--cpu-steps S]

Prints one JSON line.  `roofline` is measured live with HIP events on the
library's own stream around every launch of the dominant kernel inside the
timed region; `cpu_baseline` times the reference (oracle/_ref, single thread)
or, where it was not built, the C restatement on a bounded sample of the same
workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E

# Algorithmic bytes per unit for each step kernel ("Roofline"):
# integrate reads 17 and writes 13 doubles per body; the narrow phase and the
# candidate kernels are credited per candidate pair (bound-sphere-overlapping,
# filter-passing pair: pair write 8 + 2 x 96 B pose/geometry reads).
KERNEL_BYTES = {
    "integrate": ("body", 240.0),
    "cand_count": ("body", 48.0),
    "cand_write": ("body", 48.0),
    "cell_count": ("body", 48.0),
    "cell_fill": ("body", 48.0),
    "boundary": ("body", 48.0),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(scene, steps: int):
    """Single-threaded CPU step of the same scene: the reference itself
    (oracle/_ref) when built, else the C restatement."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import engines
    if engines.ref_available():
        w, kind = engines.RefWorld(capture=False), "reference"
    elif engines.oracle_available():
        w, kind = engines.OracleWorld(), "port"
    else:
        return None
    w.set_scene(scene)
    t0 = time.perf_counter()
    w.step(steps)
    dt = time.perf_counter() - t0
    w.close()
    return {
        "value": scene.n * steps / dt,
        "unit": "body-steps/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{steps} steps of the same {scene.n}-sphere gas, single thread, {cpu_model()}, "
                  f"nproc={os.cpu_count()}",
        "seconds": dt,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--breakdown", action="store_true", help="also time every stage (extra untimed steps)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only "
                                   "(DESIGN.md), so --gpus must be 1", "n_gpus": args.gpus}))
        sys.exit(2)

    from nativeengine_amd import scene as scn
    from nativeengine_amd.gpu import GpuWorld, KERNELS

    t0 = time.perf_counter()
    s = scn.sphere_gas(args.n, seed=0)
    log(f"scene: {s.n} spheres, L={2 * s.params.half[0]:.3f}, built in {time.perf_counter() - t0:.1f}s")

    g = GpuWorld()
    g.set_scene(s)
    g.step(max(args.warmup, 1))

    # pick the dominant stage on a few untimed steps
    stage_ms = {}
    for name in KERNELS:
        if name == "step":
            continue
        g.time_kernel(name)
        g.step(2)
        ms, k = g.kernel_time()
        stage_ms[name] = ms / max(k, 1)
    g.time_kernel("step")
    g.step(2)
    ms, k = g.kernel_time()
    step_gpu_ms = ms / max(k, 1)
    ranked = sorted(((v, k) for k, v in stage_ms.items() if k in KERNEL_BYTES), reverse=True)
    dom = ranked[0][1]
    log("stage ms/launch: " + ", ".join(f"{k}={v:.4f}" for k, v in sorted(stage_ms.items(), key=lambda t: -t[1])))
    log(f"GPU step span {step_gpu_ms:.4f} ms; dominant byte-accounted kernel: {dom}")

    # timed region: K steps, inputs resident, events on the dominant kernel
    g.time_kernel(dom)
    t0 = time.perf_counter()
    g.step(args.steps)                       # returns after the stream is synchronised
    t1 = time.perf_counter()
    dom_ms, dom_launches = g.kernel_time()
    g.time_kernel(None)
    counters = g.counters()

    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = s.n * args.steps / (t1 - t0)
    unit_kind, per_unit = KERNEL_BYTES[dom]
    units = s.n
    avg_launch_ms = dom_ms / max(dom_launches, 1)
    achieved = per_unit * units / (avg_launch_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "bytes_per_unit": per_unit,
        "units_per_launch": units,
        "avg_launch_us": round(avg_launch_ms * 1e3, 3),
    }
    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        log(f"cpu baseline: {args.cpu_steps} steps ...")
        cpu = cpu_baseline(s, args.cpu_steps)
        if cpu:
            log(f"cpu baseline {cpu['kind']}: {cpu['value']:.4g} body-steps/s ({cpu['seconds']:.1f}s)")

    out = {
        "metric": "body-steps/s at 1M bodies",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"periodic elastic sphere gas, {s.n} spheres r=0.5 m=1, L={2 * s.params.half[0]:.2f}, "
                        f"e=1 mu=0 g=0, dt=1/60, 8 iterations, seed 0",
            "bodies": s.n,
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_step_span_ms": round(step_gpu_ms, 4),
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "per_step": {"candidate_pairs": counters["candidate_pairs"], "constraints": counters["constraints"],
                     "islands": counters["islands"], "cells": counters["cells"]},
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
