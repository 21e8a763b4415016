#!/usr/bin/env python3
"""bench.py -- body-steps/s of the MI355X step at 1M bodies; batched
world-steps/s; backward/forward time ratio of differentiable rollouts.

Headline (`value`): the large terrain scene, 1,048,576 bodies in a periodic
400 x 100 x 400 domain over a 1024 x 1024-vertex value-noise heightfield:
box columns of 8 (70%), loose spheres (25%), 16-link rope chains of rigid
distance joints (5%); friction 0.6, restitution 0.1, island sleeping on,
dt 1/60, 10 iterations, FP64, seed 5.  Inputs are synthetic and resident in
HBM before the first step; a step is one full `World::step()`.  The scene's
whole 300-step run is stepped, one synchronised call per step, every call
timed.  `value` is the body-steps/s of --steps K of those steps spread evenly
over the run after its first --warmup W steps (all 300 by default);
`whole_run` is the whole run's.  `cpu_baseline` times the C restatement on
one host core for one step at steps 10/100/200/300 of the 65,536-body
instance of the same scene, from the GPU's state there, and
`speedup_vs_cpu` compares the GPU's own steps 10/100/200/300.

Other lines in the same JSON object:
  gas      the 1,048,576-sphere periodic elastic gas (r 0.5, L 253.98, e 1,
           mu 0, no gravity, 8 iterations, seed 0), the plain-method 1M point;
  gas4096  the 4096-sphere gas of side 40 over 1000 steps, with its momentum
           and energy drift beside the reference's own (golden fixture);
  pile     the 65,536-body mixed-shape pile (40% boxes, 30% spheres, 20%
           capsules, 10% hulls, sizes as specified) on a ground box inside
           100 x 200 x 100 walls, 10 iterations, its whole 600-step run, with
           the C restatement at steps 10/150/300/450/600; at the end of its
           run, scene queries on the settled pile ("queries": 1M rays cast
           down over its floor, the same as swept spheres, raycastAll,
           65,536 overlapSphere queries, the C restatement on a sample);
  batched  1024 periodic worlds of log-uniform size in [64, 4096] stepped in
           one set of launches (HIP-graph replay, autotuned) over their whole
           300-step run, against looping the plain single-world step over a
           sample of the same worlds at three points of the run;
  adjoint  256 worlds of 256 spheres + 256 boxes on a ground box, 120 steps,
           checkpoint every 16: backward time / forward time.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-large] [--no-gas] [--no-pile]
                    [--no-queries] [--no-config1] [--no-batched] [--no-adjoint] [--traffic-probe]

Prints one JSON line.  `roofline` is the dominant stage of the run: every
stage is timed with HIP events on the library's own stream in every 9th
step of the run, in rotation, and credited with the algorithmic bytes of
those steps; `traffic` is the HBM bytes per launch of that stage from the
committed rocprofv3 PMC summary.
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

# Algorithmic bytes per unit for each step kernel:
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


def cpu_baseline(scene, steps: int, what: str):
    """Single-threaded CPU step of the same scene: the reference itself
    (oracle/_ref) when built and able to express the scene (a cubic box,
    8 iterations), else the C restatement."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import engines
    w = kind = None
    if engines.ref_available():
        w, kind = engines.RefWorld(capture=False), "reference"
        try:
            w.set_scene(scene)
        except RuntimeError:        # not expressible by the reference
            w.close()
            w = None
    if w is None:
        if not engines.oracle_available():
            return None
        w, kind = engines.OracleWorld(), "port"
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
        "sample": f"{steps} steps of {what}, single thread, {cpu_model()}, nproc={os.cpu_count()}",
        "seconds": dt,
    }


# Kernels of each stage, by name (rocprofv3 names them with their namespace).
STAGE_OF_KERNEL = (("narrow", ("k_narrow", "k_epa", "k_mesh", "k_snap_counter")),
                   ("solve", ("k_wf_", "k_solve_", "k_copy_cache")))


def _kernel_base(name: str) -> str:
    """`(anonymous namespace)::k_epa_lds<8, 12, 8>(...)` -> `k_epa_lds<8, 12, 8>`."""
    import re
    name = re.sub(r"\(anonymous namespace\)::|nea::|^void ", "", name)
    depth, out = 0, []
    for ch in name:   # cut the argument list, keep template arguments
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out).strip()


def _stage_of(kernel: str):
    for stage, prefixes in STAGE_OF_KERNEL:
        if kernel.startswith(prefixes):
            return stage
    return None


def _counter_sums(path: str, counter: str):
    """Per kernel: (sum of the counter over its dispatches, dispatches), and
    the number of k_integrate dispatches (one per step)."""
    import csv
    sums, calls = {}, {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            if row.get("Counter_Name") != counter:
                continue
            k = _kernel_base(row["Kernel_Name"])
            sums[k] = sums.get(k, 0.0) + float(row["Counter_Value"])
            calls[k] = calls.get(k, 0) + 1
    steps = sum(v for k, v in calls.items() if k.startswith("k_integrate"))
    return sums, steps


def pmc_summary(probe_f: str, probe_w: str, run_f: str, run_w: str) -> str:
    """HBM bytes per step of each stage and kernel: FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch) summed over the run, each scaled so that the probe
    kernel's reading equals the bytes it moves (reads 512 MiB, writes 256 MiB),
    divided by the steps (k_integrate dispatches)."""
    pf, _ = _counter_sums(probe_f, "FETCH_SIZE")
    pw, _ = _counter_sums(probe_w, "WRITE_SIZE")
    kf = (2 * 8 << 25) / (pf.get("k_fp_probe", 0.0) * 1024.0)
    kw = (8 << 25) / (pw.get("k_fp_probe", 0.0) * 1024.0)
    rf, steps = _counter_sums(run_f, "FETCH_SIZE")
    rw, steps_w = _counter_sums(run_w, "WRITE_SIZE")
    steps = max(steps, 1)
    per = {}
    for k in set(rf) | set(rw):
        per[k] = (rf.get(k, 0.0) * 1024.0 * kf / steps, rw.get(k, 0.0) * 1024.0 * kw / max(steps_w, 1))
    lines = [f"probe (`bench.py --traffic-probe`: k_fp_probe reads 512 MiB, writes 256 MiB): FETCH_SIZE scale {kf:.4f}, "
             f"WRITE_SIZE scale {kw:.4f}", "", f"steps in the run (k_integrate dispatches): {steps}", ""]
    stage_tot = {}
    for k, (r, w) in per.items():
        st = _stage_of(k)
        if st:
            a = stage_tot.setdefault(st, [0.0, 0.0])
            a[0] += r
            a[1] += w
    for st, (r, w) in sorted(stage_tot.items()):
        lines.append(f"- {st}: read {r / 1e9:.2f} GB, write {w / 1e9:.2f} GB, total {(r + w) / 1e9:.2f} GB per step")
    lines += ["", "top kernels (GB per step, scaled):"]
    for k, (r, w) in sorted(per.items(), key=lambda kv: -(kv[1][0] + kv[1][1]))[:16]:
        lines.append(f"- `{k}`: read {r / 1e9:.2f}, write {w / 1e9:.2f}")
    lines += ["", "traffic_bytes_per_launch: " + " ".join(f"{st}={int(r + w)}" for st, (r, w) in sorted(stage_tot.items()))]
    return "\n".join(lines)


def kernel_summary(path: str, top: int = 30) -> str:
    """The per-kernel table of a rocprofv3 --stats kernel_stats.csv."""
    import csv
    rows = []
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            rows.append((_kernel_base(row["Name"]), int(row["Calls"]), float(row["TotalDurationNs"]),
                         float(row["AverageNs"]), float(row["Percentage"])))
    total = sum(r[2] for r in rows)
    steps = sum(r[1] for r in rows if r[0].startswith("k_integrate"))
    out = [f"total kernel time {total / 1e6:.1f} ms over {steps} steps (k_integrate dispatches): "
           f"{total / 1e6 / max(steps, 1):.2f} ms per step", "",
           "| kernel | calls | total ms | ms per step | avg us | % |", "|---|---|---|---|---|---|"]
    for k, c, t, a, p in sorted(rows, key=lambda r: -r[2])[:top]:
        out.append(f"| `{k}` | {c} | {t / 1e6:.1f} | {t / 1e6 / max(steps, 1):.3f} | {a / 1e3:.1f} | {p:.2f} |")
    return "\n".join(out)


def measured_traffic(stage: str, profile: str = "large_terrain_1M_traffic.md"):
    """HBM bytes per launch of a stage from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by a kernel of known traffic), or
    None.  PMC counters need their own profiler runs, so bench.py does not
    measure them itself."""
    path = os.path.join(ROOT, "profiles", profile)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    for item in line.split(":", 1)[1].split():
                        k, v = item.split("=")
                        if k == stage:
                            return float(v)
    except OSError:
        pass
    return None


def stage_bytes(name: str, iterations: int):
    """Algorithmic bytes per unit of a step stage: {unit kind: bytes}."""
    table = {
        "integrate": {"body": 240.0},
        "cell_count": {"body": 48.0},
        "cell_fill": {"body": 48.0},
        "cand_count": {"body": 48.0},
        "cand_write": {"body": 48.0},
        "boundary": {"body": 48.0},
        "narrow": {"pair": 200.0, "mesh_triangle": 72.0},
        "cons_build": {"contact": 352.0},
        "solve": {"contact": 440.0 * (iterations + 1)},
    }
    return table.get(name)


STAGES = ["integrate", "cell_count", "cell_fill", "cand_count", "cand_write", "narrow", "cons_build", "solve",
          "boundary"]


def timed_steps(run: int, warmup: int, k: int):
    """The steps (1-based) whose times make `value`: k steps spread evenly
    over the run after its first `warmup` steps (midpoints of k equal
    strata), or all of them when k covers the rest."""
    rest = run - warmup
    if k >= rest:
        return list(range(warmup + 1, run + 1))
    return sorted({warmup + 1 + int((i + 0.5) * rest / k) for i in range(k)})


def scene_run(s, run: int, warmup: int, k: int, label: str, marks=(), after=None):
    """One resident scene stepped through its whole run of `run` single-step
    calls, each bracketed by a stream synchronisation (the call returns after
    it) and a read of the step's counters.  Every step's wall time is kept;
    `value` comes from the k timed steps spread over the run.  Stage times
    come from HIP events on the library's stream, one stage per step in
    rotation over the whole run, with that step's algorithmic units, so the
    dominant stage and its roofline are measured over the same run.  At the
    steps in `marks` the GPU's state is saved for the CPU's phase-matched
    samples.  `after(g)`, if given, runs on the world at the end of the run
    and its result is kept under "queries"."""
    import tempfile
    import numpy as np
    from nativeengine_amd.gpu import GpuWorld
    iters = int(s.params.solver_iterations)
    t0 = time.perf_counter()
    g = GpuWorld(strict=False)
    g.set_scene(s)
    log(f"{label}: upload {time.perf_counter() - t0:.1f}s")
    timed = set(timed_steps(run, warmup, k))
    step_ms = []
    stage = {name: {"ms": 0.0, "steps": 0, "units": {"body": 0, "pair": 0, "mesh_triangle": 0, "contact": 0}}
             for name in STAGES}
    units_all = {"body": 0, "pair": 0, "mesh_triangle": 0, "contact": 0}
    capped_steps = 0
    tmp = tempfile.mkdtemp(prefix="bench_states_")
    states = {}
    tstart = last_log = time.perf_counter()
    for t in range(1, run + 1):
        name = STAGES[(t - 1) % len(STAGES)]
        g.time_kernel(name)
        calls0 = g.epa_capped_calls
        ts = time.perf_counter()
        g.step(1)
        c = g.counters()
        now = time.perf_counter()
        step_ms.append((now - ts) * 1e3)
        if g.epa_capped_calls != calls0:
            capped_steps += 1
        ms, launches = g.kernel_time()
        u = {"body": s.n, "pair": c["candidate_pairs"], "mesh_triangle": c["mesh_triangles"],
             "contact": c["constraints"]}
        for kk in units_all:
            units_all[kk] += u[kk]
        if launches:
            st = stage[name]
            st["ms"] += ms
            st["steps"] += 1
            for kk in st["units"]:
                st["units"][kk] += u[kk]
        if t in marks:
            path = os.path.join(tmp, f"state_{t}.npz")
            _save_state(g, s, path)
            states[t] = path
        if now - last_log > 10.0:
            log(f"{label}: step {t}/{run}, {(now - tstart) * 1e3 / t:.2f} ms/step so far, last {step_ms[-1]:.2f} ms, "
                f"{c['constraints']} constraints, largest island {c['max_island']}")
            last_log = now
    g.time_kernel(None)
    last = g.counters()
    reruns, trunc = g.epa_stats()
    epa_queues = g.epa_queues()
    fl = g.body_aux()[0] if s.params.sleep_enabled else None
    after_out = after(g) if after else None
    g.close()
    sel = sorted(timed)
    timed_ms = [step_ms[t - 1] for t in sel]
    value = s.n * len(sel) / (sum(timed_ms) * 1e-3)
    mean_ms = {kk: v["ms"] / max(v["steps"], 1) for kk, v in stage.items()}
    ranked = sorted(((v, kk) for kk, v in mean_ms.items() if stage_bytes(kk, iters)), reverse=True)
    dom = ranked[0][1]
    st = stage[dom]
    per = stage_bytes(dom, iters)
    bytes_total = sum(per[kk] * st["units"][kk] for kk in per)
    avg_launch_ms = st["ms"] / max(st["steps"], 1)
    achieved = bytes_total / (st["ms"] * 1e-3) / 1e9 if st["ms"] > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": None,
        "bytes_per_unit": per,
        "units_per_launch": {kk: round(st["units"][kk] / max(st["steps"], 1), 1) for kk in per},
        "avg_launch_us": round(avg_launch_ms * 1e3, 3),
        "launches_timed": st["steps"],
        "window": f"the stage's span on the library's stream in every {len(STAGES)}th step of the {run}-step run "
                  f"({st['steps']} steps), one launch = the stage's span in one step",
    }
    pct = np.percentile(step_ms, [50, 90, 100])
    out = {
        "value": value,
        "unit": "body-steps/s",
        "run_steps": run,
        "steps": len(sel),
        "warmup": warmup,
        "timed_steps": sel if len(sel) < 40 else f"steps {sel[0]}..{sel[-1]} (all)",
        "ms_per_step": sum(timed_ms) / len(sel),
        "whole_run": {"value": s.n * run / (sum(step_ms) * 1e-3), "ms_per_step": sum(step_ms) / run,
                      "step_ms_p50_p90_max": [round(float(x), 3) for x in pct]},
        "bodies": s.n,
        "roofline": roofline,
        "stage_ms_mean": {kk: round(v, 4) for kk, v in mean_ms.items()},
        "per_step_mean": {"candidate_pairs": units_all["pair"] / run, "mesh_triangles": units_all["mesh_triangle"] / run,
                          "constraints": units_all["contact"] / run},
        "last_step": {"candidate_pairs": last["candidate_pairs"], "constraints": last["constraints"],
                      "islands": last["islands"], "largest_island": last["max_island"]},
        "epa_large_workspace_reruns_last_step": reruns,
        "epa_truncated_last_step": trunc,
        "epa_tier_queues_last_step": epa_queues,
        "steps_with_epa_at_face_cap": capped_steps,
        "step_ms_at": {t: round(step_ms[t - 1], 3) for t in marks},
    }
    if fl is not None:
        out["sleeping_bodies_at_end"] = int(np.sum((fl & 8) != 0))
    if after_out is not None:
        out["queries"] = after_out
    return out, states, tmp


RAY_BYTES = 128   # per ray: origin, direction, search distance, cast radius, mask in; body, distance, point, normal out


def query_bench(g, s, no_cpu: bool):
    """Scene queries on the settled pile: a 1024 x 1024 grid of rays cast
    down over the box's whole floor (a depth camera; nearest hit), the same
    grid as swept spheres of radius 0.25, raycastAll on its 262,144 middle
    rays, and 65,536 overlapSphere queries in the occupied layer; the ray
    kernel's time from HIP events on the library's stream.  The CPU baseline
    is the C restatement on 512 of the rays and 256 of the spheres at the
    same state."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    h = np.asarray(s.params.half, float)
    k = 1024
    gx, gz = np.meshgrid(np.linspace(-h[0] + 0.5, h[0] - 0.5, k), np.linspace(-h[2] + 0.5, h[2] - 0.5, k))
    org = np.column_stack([gx.ravel(), np.full(k * k, h[1] - 0.5), gz.ravel()])
    rng = np.random.default_rng(0)
    d = np.column_stack([rng.normal(0, 0.05, k * k), -np.ones(k * k), rng.normal(0, 0.05, k * k)])
    md = 2.0 * h[1]
    out = {"rays": k * k}
    for label, fn in (("raycast", lambda: g.raycast(org, d, md)), ("sphere_cast", lambda: g.sphere_cast(org, 0.25, d, md))):
        fn()
        g.time_kernel("ray")
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            res = fn()
        wall = (time.perf_counter() - t0) / reps
        ms, launches = g.kernel_time()
        kms = ms / max(launches, 1)
        out[label] = {"kernel_ms": round(kms, 3), "call_ms": round(wall * 1e3, 3),
                      "rays_per_s": k * k / (kms * 1e-3), "rays_per_s_call": k * k / wall,
                      "hit_fraction": round(float((res["body"] >= 0).mean()), 4)}
    g.time_kernel(None)
    kms = out["raycast"]["kernel_ms"]
    achieved = RAY_BYTES * k * k / (kms * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "kernel": "k_ray", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                       "bytes_per_ray": RAY_BYTES,
                       "note": "the fold is bound by its FP64 ray-body tests over LDS-staged tiles, not by HBM: "
                               "the fraction only prices the rays' own bytes"}
    mid = (np.abs(gx.ravel()) < h[0] / 2) & (np.abs(gz.ravel()) < h[2] / 2)
    t0 = time.perf_counter()
    allh = g.raycast_all(org[mid], d[mid], md)
    out["raycast_all"] = {"rays": int(mid.sum()), "call_ms": round((time.perf_counter() - t0) * 1e3, 3),
                          "hits": int(len(allh["body"]))}
    x = g.state()[0]
    dyn = s.inv_mass > 0
    lo, hi = np.percentile(x[dyn], 1, axis=0), np.percentile(x[dyn], 99, axis=0)
    nq = 65536
    cs = rng.uniform(lo, hi, (nq, 3))
    rs = rng.uniform(0.3, 1.5, nq)
    from nativeengine_amd import queries as Q
    shapes = Q.spheres(cs, rs)
    g.overlap(shapes[:1024])
    t0 = time.perf_counter()
    ov = g.overlap(shapes)
    wall = time.perf_counter() - t0
    out["overlap_sphere"] = {"queries": nq, "call_ms": round(wall * 1e3, 3), "queries_per_s": nq / wall,
                             "hits": int(len(ov["body"]))}
    if not no_cpu:
        import engines
        o = engines.OracleWorld()
        o.set_scene(s)
        xs, vs, qs, ws = g.state()
        o.set_state(xs, vs, qs, ws)
        sel = rng.choice(k * k, 512, replace=False)
        t0 = time.perf_counter()
        cr = o.raycast(org[sel], d[sel], md)
        tr = time.perf_counter() - t0
        same = bool(all(np.array_equal(cr[f], g.raycast(org[sel], d[sel], md)[f]) for f in cr))
        t0 = time.perf_counter()
        o.overlap(shapes[:256])
        to = time.perf_counter() - t0
        out["cpu_baseline"] = {"raycast_rays_per_s": 512 / tr, "overlap_queries_per_s": 256 / to, "cores": 1,
                               "kind": "port", "sample": "512 of the rays and 256 of the spheres, same state",
                               "sample_equal_to_gpu": same}
        out["speedup_vs_cpu"] = {"raycast": round(out["raycast"]["rays_per_s_call"] / (512 / tr), 1),
                                 "overlap_sphere": round(out["overlap_sphere"]["queries_per_s"] / (256 / to), 1)}
    log(f"queries: raycast {out['raycast']['kernel_ms']} ms kernel for {k * k} rays, "
        f"overlap {out['overlap_sphere']['call_ms']} ms for {nq} spheres")
    return out


def _save_state(g, s, path):
    """The world's bodies as the GPU holds them now (pose, velocity, sleep
    state) with the scene's shapes, materials, joints and meshes."""
    import numpy as np
    x, v, q, w = g.state()
    fl, tm, im, ii = g.body_aux()
    ims, iis = g.sleep_store()
    np.savez(path, x=x, v=v, q=q, w=w, flags=fl, timer=tm, im=im, ii=ii, ims=ims, iis=iis)


def _build_scene(spec: str):
    """A bench scene from its spec string "large,<n>,<seed>" or "pile,<n>,<seed>"."""
    from nativeengine_amd import scene as scn
    kind, n, seed = spec.split(",")
    if kind == "large":
        return scn.large_scene(int(n), seed=int(seed))
    return pile_scene(int(n), int(seed))


def pile_scene(n: int, seed: int = 0):
    """The mixed-shape pile as specified: n dynamic bodies on a static ground
    box inside 100 x 200 x 100 reflective walls, 10 iterations."""
    from nativeengine_amd import scene as scn
    return scn.mixed_pile(n, half=(50.0, 100.0, 50.0), seed=seed, iterations=10, sizes="config")


def cpu_sample_main(state_path, spec):
    """Child process: rebuild the scene, load the saved state, time one step
    of the C restatement."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import engines
    s = _build_scene(spec)
    d = np.load(state_path)
    s.x, s.v, s.q, s.w = d["x"], d["v"], d["q"], d["w"]
    s.flags = d["flags"].astype(np.uint32)
    s.sleep_timer, s.inv_mass, s.inv_inertia = d["timer"], d["im"], d["ii"]
    s.inv_mass_store, s.inv_inertia_store = d["ims"], d["iis"]
    o = engines.OracleWorld()
    o.set_scene(s)
    t0 = time.perf_counter()
    o.step(1)
    print(time.perf_counter() - t0)


def _cpu_one_step(spec, state_path, timeout):
    """Seconds the C restatement takes for one step from a saved state, in a
    child process (a single step can meet one of the reference's
    non-converging EPA runs, which take minutes): None past `timeout`."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-sample", state_path, "--cpu-scene", spec]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return None
    if r.returncode != 0:
        log(r.stderr[-2000:])
        return None
    return float(r.stdout.strip().splitlines()[-1])


def phase_cpu(spec: str, n: int, states: dict, timeout: float, what: str):
    """The C restatement (one thread) for one step from each saved GPU state
    (the GPU's states are the restatement's, bit for bit)."""
    times = {}
    for m in sorted(states):
        sec = _cpu_one_step(spec, states[m], timeout)
        log(f"cpu ({what}): step {m}: " + (f"{sec:.2f}s" if sec is not None else f"> {timeout:.0f}s (stopped)"))
        times[m] = sec
    # a step that did not finish counts as the timeout (the CPU figure is then an upper bound)
    total = sum(t if t is not None else timeout for t in times.values())
    return {
        "value": n * len(times) / total,
        "unit": "body-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"one step at each of steps {sorted(times)} of {what}, each from the GPU's state at that step, "
                  f"single thread, {cpu_model()}, nproc={os.cpu_count()}",
        "seconds": total,
        "seconds_at": {m: (round(t, 3) if t is not None else None) for m, t in times.items()},
        "steps_over_timeout": sum(1 for t in times.values() if t is None),
        "upper_bound": any(t is None for t in times.values()),
    }


def _phase_matched(r, cpu, n_gpu, n_cpu):
    """GPU over CPU body-steps/s on the same steps of the run."""
    gpu_ms = sum(r["step_ms_at"][m] for m in cpu["seconds_at"])
    gpu_rate = n_gpu * len(cpu["seconds_at"]) / (gpu_ms * 1e-3)
    return {"gpu_value": gpu_rate, "cpu_value": cpu["value"], "ratio": round(gpu_rate / cpu["value"], 1),
            "steps": sorted(cpu["seconds_at"]), "gpu_bodies": n_gpu, "cpu_bodies": n_cpu}


def large_bench(args):
    """The large terrain scene at 1,048,576 bodies: the headline line."""
    from nativeengine_amd import scene as scn
    from nativeengine_amd.gpu import GpuWorld
    t0 = time.perf_counter()
    s = scn.large_scene(args.large_n, seed=5)
    h = s.params.half
    log(f"large scene: {s.n} bodies, {len(s.joints)} joints, {len(s.meshes[0]['tris'])} triangles, "
        f"built in {time.perf_counter() - t0:.1f}s")
    marks = [10, 100, 200, 300]
    r, _, _ = scene_run(s, 300, args.warmup, args.steps, "large", marks=marks)
    r["traffic_profile"] = "profiles/large_terrain_1M_traffic.md"
    r["roofline"]["traffic"] = measured_traffic(r["roofline"]["kernel"], "large_terrain_1M_traffic.md")
    if not args.no_cpu:
        # the CPU's phase samples on the 65,536-body instance of the same scene
        small = scn.large_scene(args.cpu_n, seed=5)
        g = GpuWorld(strict=False)
        g.set_scene(small)
        states, done = {}, 0
        import tempfile
        tmp = tempfile.mkdtemp(prefix="bench_cpu_")
        for m in marks:
            g.step(m - done)
            done = m
            path = os.path.join(tmp, f"state_{m}.npz")
            _save_state(g, small, path)
            states[m] = path
        g.close()
        cpu = phase_cpu(f"large,{args.cpu_n},5", args.cpu_n, states, args.cpu_timeout,
                        f"the {args.cpu_n}-body instance of the same scene (same density, side "
                        f"{400.0 * (args.cpu_n / float(1 << 20)) ** 0.5:.0f})")
        r["cpu_baseline"] = cpu
        r["phase_matched"] = _phase_matched(r, cpu, s.n, args.cpu_n)
        r["speedup_vs_cpu"] = r["phase_matched"]["ratio"]
        log(f"large: cpu {cpu['value']:.4g} body-steps/s; GPU/CPU at the same steps {r['speedup_vs_cpu']}")
    else:
        r["cpu_baseline"], r["speedup_vs_cpu"] = None, None
    r["workload"] = (f"large terrain scene: {s.n} bodies (mesh + {int((s.shape == 2).sum())} boxes in columns of 8 + "
                     f"{int((s.shape == 0).sum())} spheres, {len(s.joints)} rope joints in 16-link ropes), periodic "
                     f"{2 * h[0]:.0f}x{2 * h[1]:.0f}x{2 * h[2]:.0f}, {len(s.meshes[0]['verts'])}-vertex value-noise "
                     f"heightfield amplitude 8, mu 0.6, e 0.1, sleeping on, dt 1/60, "
                     f"{s.params.solver_iterations} iterations, 300 steps, seed 5")
    return r


def pile_bench(args):
    """The mixed-shape pile under gravity (65,536 bodies, reflective walls),
    its whole 600-step run, with the C restatement at five steps of it."""
    t0 = time.perf_counter()
    s = pile_scene(args.pile_n, 0)
    log(f"pile: {s.n} bodies, built in {time.perf_counter() - t0:.1f}s")
    marks = [10, 150, 300, 450, 600]
    after = None if args.no_queries else (lambda g: query_bench(g, s, args.no_cpu))
    r, states, tmp = scene_run(s, 600, 0, args.pile_steps, "pile", marks=marks if not args.no_cpu else (), after=after)
    r["traffic_profile"] = "profiles/mixed_pile_65536_traffic.md"
    r["roofline"]["traffic"] = measured_traffic(r["roofline"]["kernel"], "mixed_pile_65536_traffic.md")
    if not args.no_cpu:
        cpu = phase_cpu(f"pile,{args.pile_n},0", s.n, states, args.cpu_timeout, "the same pile")
        r["cpu_baseline"] = cpu
        r["phase_matched"] = _phase_matched(r, cpu, s.n, s.n)
        r["speedup_vs_cpu"] = r["phase_matched"]["ratio"]
        log(f"pile: cpu {cpu['value']:.4g} body-steps/s; GPU/CPU at the same steps {r['speedup_vs_cpu']}")
    r["workload"] = (f"mixed-shape pile: {s.n - 1} dynamic bodies (40% boxes half extents U[0.2,0.6]^3, 30% spheres "
                     f"r U[0.2,0.6], 20% capsules r 0.2 half-height 0.4, 10% 12-vertex hulls r U[0.2,0.6]) on a "
                     f"jittered lattice of spacing 2.3 with random orientations, on a static ground box, reflective "
                     f"100x200x100 box, g -9.81, mu 0.5, e 0.2, dt 1/60, 10 iterations, warm start, 600 steps, seed 0")
    return r


def config1_bench(args):
    """The elastic gas as specified for the CPU reference: 4096 spheres,
    side 40, 1000 steps; ms per step and the conservation of momentum and
    energy, beside the reference's own drift from its golden fixture."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import fixtures
    from nativeengine_amd.gpu import GpuWorld
    s, steps, out = fixtures.load_scene("gas4096")
    g = GpuWorld()
    g.set_scene(s)
    g.use_graph(True)       # each step replays its captured graph (0.149 -> 0.142 ms a step measured)
    p0, L0, e0 = g.diagnostics()
    g.step(1)
    g.set_scene(s)          # the first call loaded the kernels; start again from the initial state
    g.reset_warm_cache()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.step(1)
    t1 = time.perf_counter()
    p1, L1, e1 = g.diagnostics()
    x = g.state()[0]
    g.close()
    ref = out["diag"]
    same = bool(np.array_equal(np.concatenate([p1, L1, [e1]]), ref[-1]) and np.array_equal(x, out["x"]))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(s, 200, "the same gas (first 200 of its 1000 steps)")
    return {
        "value": s.n * steps / (t1 - t0),
        "unit": "body-steps/s",
        "ms_per_step": (t1 - t0) * 1e3 / steps,
        "steps": steps,
        "bodies": s.n,
        "momentum_drift": float(np.max(np.abs(p1 - p0))),
        "energy_drift_rel": float((e1 - e0) / e0),
        "reference_energy_drift_rel": float((ref[-1][6] - ref[0][6]) / ref[0][6]),
        "reference_momentum_drift": float(np.max(np.abs(ref[-1][:3] - ref[0][:3]))),
        "bitwise_equal_to_reference": same,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(s.n * steps / (t1 - t0) / cpu["value"], 1) if cpu else None,
        "note": "one synchronised call per step, replayed from a HIP graph; 4096 bodies leave the GPU "
                "latency-bound, so no roofline is given",
        "workload": "periodic elastic sphere gas: 4096 spheres r 0.5 m 1, side 40, rejection-sampled (seed 0), "
                    "v N(0,1) with the net momentum removed, e 1, mu 0, g 0, dt 1/60, 8 iterations, 1000 steps",
    }


def _loop_sample(g, b, sample, loop_steps):
    """ms per world-step of the plain single-world step looped over the
    sampled worlds, each in its own context from the batch's current state
    and warm cache."""
    import numpy as np
    from nativeengine_amd.gpu import GpuWorld
    x, v, q, w = g.state()
    keys, imp = g.export_warm_cache()
    ki, kj = (keys >> np.uint64(32)).astype(np.int64), (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    ctxs = []
    for wi in sample:
        a, e = int(b.offsets[wi]), int(b.offsets[wi + 1])
        ws = b.world(int(wi))
        ws.x, ws.v, ws.q, ws.w = x[a:e].copy(), v[a:e].copy(), q[a:e].copy(), w[a:e].copy()
        c = GpuWorld(strict=False)
        c.set_scene(ws)
        sel = (ki >= a) & (ki < e)
        c.import_warm_cache(((ki[sel] - a).astype(np.uint64) << np.uint64(32)) | (kj[sel] - a).astype(np.uint64),
                            imp[sel])
        ctxs.append(c)
    for c in ctxs:       # first call of each context (launch sizing) outside the timing
        c.step(1)
    t0 = time.perf_counter()
    for _ in range(loop_steps):
        for c in ctxs:
            c.step(1)
    t1 = time.perf_counter()
    for c in ctxs:
        c.close()
    return (t1 - t0) * 1e3 / (loop_steps * len(ctxs))


BATCH_CALL_STEPS = 10


def batched_bench(worlds: int, loop_worlds: int, loop_steps: int, graph: bool = True):
    """Batched world-steps/s over the whole 300-step run of the batch (HIP
    graph replay; the autotuner runs after step 30, from a snapshot it
    restores, outside the timing), against looping the plain single-world
    step over a sample of the same worlds at steps 30, 150 and 300."""
    import numpy as np
    from nativeengine_amd import scene as scn
    from nativeengine_amd.gpu import GpuWorld
    t0 = time.perf_counter()
    b = scn.batched_worlds(worlds, seed=1)
    log(f"batched: {b.nworlds} worlds, {b.scene.n} bodies, built in {time.perf_counter() - t0:.1f}s")
    g = GpuWorld(strict=False)
    g.set_scene(b.scene)
    g.set_worlds(b.offsets, b.half, b.restitution, b.friction)
    g.use_graph(True)
    rng = np.random.default_rng(7)
    sample = np.sort(rng.choice(b.nworlds, size=min(loop_worlds, b.nworlds), replace=False))
    timed_s, loop_ms, capped, tune = 0.0, {}, 0, None
    done = 0
    for chunk_end in (30, 150, 300):
        calls0 = g.epa_capped_calls
        t0 = time.perf_counter()
        # calls of BATCH_CALL_STEPS graph-replayed steps: a call whose step
        # outgrows a buffer reruns from the call's start with the buffer grown
        while done < chunk_end:
            k = min(BATCH_CALL_STEPS, chunk_end - done)
            g.step(k)
            done += k
        timed_s += time.perf_counter() - t0
        capped += g.epa_capped_calls - calls0
        if tune is None:
            tune = g.autotune(steps=3)
            log(f"autotuner: {tune}")
            if not graph:   # kernel by kernel (what a kernel trace profiles), whatever the autotuner chose
                g.use_graph(False)
        loop_ms[chunk_end] = _loop_sample(g, b, sample, loop_steps)
        log(f"batched: step {done}, {timed_s * 1e3 / done:.2f} ms/step so far; loop {loop_ms[chunk_end]:.3f} ms "
            f"per world-step")
    reruns, trunc = g.epa_stats()
    g.close()
    batched_ms = timed_s * 1e3 / 300
    batched_wps = b.nworlds / (batched_ms * 1e-3)
    loop_wps = 1e3 / (sum(loop_ms.values()) / len(loop_ms))
    return {
        "metric": "batched world-steps/s",
        "value": batched_wps,
        "unit": "world-steps/s",
        "ms_per_step": batched_ms,
        "steps": 300,
        "worlds": b.nworlds,
        "bodies": int(b.scene.n),
        "body_steps_per_s": b.scene.n / (batched_ms * 1e-3),
        "loop_world_steps_per_s": loop_wps,
        "loop_ms_per_world_step_at": {k: round(v, 4) for k, v in loop_ms.items()},
        "loop_sample": f"{len(sample)} of the worlds, each in its own context from the batch's state and warm cache "
                       f"at steps 30, 150 and 300, {loop_steps} steps each",
        "vs_loop": batched_wps / loop_wps,
        "autotune": tune,
        "graph_replay": bool(graph and tune and tune.get("graph")),
        "calls_with_epa_at_face_cap": capped,
        "epa_large_workspace_reruns_last_call": reruns,
        "epa_truncated_last_call": trunc,
        "config": "1024 periodic worlds, N_w log-uniform [64,4096] seed 1, half spheres r U[0.2,0.6] / half boxes "
                  "he U[0.2,0.6]^3, 20% volume fraction, mu_w U[0,0.8], e_w U[0,1], g=0, dt=1/60, 8 iterations, "
                  "300 steps",
    }


def adjoint_bench(worlds: int, steps: int, every: int):
    """Forward rollout and adjoint times on the differentiable-rollout workload."""
    import numpy as np
    from nativeengine_amd import scene as scn
    from nativeengine_amd.gpu import GpuWorld
    b = scn.adjoint_worlds(worlds, n_spheres=256, n_boxes=256, side=16.0, seed=2)
    nw = b.nworlds
    g = GpuWorld(strict=False)
    g.set_scene(b.scene)
    g.set_worlds(b.offsets, b.half, b.restitution, b.friction)
    target = g.rollout(b.scene.v, np.full(nw, 0.4), np.full(nw, 0.3), steps, every)
    mu, e = np.full(nw, 0.5), np.full(nw, 0.2)
    fwd, bwd = [], []
    for _ in range(2):
        g.set_state(b.scene.x, b.scene.v, b.scene.q, b.scene.w)
        t0 = time.perf_counter()
        x = g.rollout(b.scene.v, mu, e, steps, every)
        t1 = time.perf_counter()
        gv, gmu, ge = g.rollout_backward(2.0 * (x - target) / x.size, nw)
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        bwd.append(t2 - t1)
    g.close()
    f, bk = min(fwd), min(bwd)
    return {
        "metric": "backward/forward time ratio",
        "value": bk / f,
        "unit": "ratio",
        "higher_is_better": False,
        "forward_s": f,
        "backward_s": bk,
        "worlds": nw,
        "bodies": int(b.scene.n),
        "steps": steps,
        "checkpoint_every": every,
        "body_steps_per_s_forward": b.scene.n * steps / f,
        "loss": float(np.mean((x - target) ** 2)),
        "grad_norms": {"v0": float(np.linalg.norm(gv)), "mu": float(np.linalg.norm(gmu)), "e": float(np.linalg.norm(ge))},
        "gradient": "exact adjoint of the step with each step's contact set fixed (integrate, constraint build, "
                    "warm start, 8 Gauss-Seidel iterations); checked against finite differences of the oracle's "
                    "fixed-contact replay in tests/test_gpu_adjoint.py",
        "config": "256 worlds x (256 spheres r=0.4 + 256 boxes he=0.3 + static ground box), periodic side 16, "
                  "g=-9.81, dt=1/60, 8 iterations, 120 steps, loss MSE vs targets at mu 0.4 / e 0.3, "
                  "evaluated at mu 0.5 / e 0.2",
    }


def gas_bench(args):
    """The 1M-sphere gas: the headline body-steps/s line, with roofline and
    CPU baseline."""
    from nativeengine_amd import scene as scn
    from nativeengine_amd.gpu import GpuWorld, KERNELS

    t0 = time.perf_counter()
    s = scn.sphere_gas(args.n, seed=0)
    log(f"scene: {s.n} spheres, L={2 * s.params.half[0]:.3f}, built in {time.perf_counter() - t0:.1f}s")

    g = GpuWorld()
    g.set_scene(s)
    g.step(max(args.gas_warmup, 1))

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
    g.step(args.gas_steps)                   # returns after the stream is synchronised
    t1 = time.perf_counter()
    dom_ms, dom_launches = g.kernel_time()
    g.time_kernel(None)
    counters = g.counters()

    ms_per_step = (t1 - t0) * 1e3 / args.gas_steps
    value = s.n * args.gas_steps / (t1 - t0)
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
    g.close()
    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        log(f"cpu baseline: {args.cpu_steps} steps ...")
        cpu = cpu_baseline(s, args.cpu_steps, f"the same {s.n}-sphere gas")
        if cpu:
            log(f"cpu baseline {cpu['kind']}: {cpu['value']:.4g} body-steps/s ({cpu['seconds']:.1f}s)")

    return {
        "metric": "body-steps/s at 1M bodies; batched world-steps/s; backward/forward time ratio",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.gas_steps,
        "warmup": args.gas_warmup,
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
        "speedup_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "gpu_step_span_ms": round(step_gpu_ms, 4),
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "per_step": {"candidate_pairs": counters["candidate_pairs"], "constraints": counters["constraints"],
                     "islands": counters["islands"], "cells": counters["cells"]},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300,
                    help="timed steps of the large terrain scene's 300-step run, spread evenly over it")
    ap.add_argument("--warmup", type=int, default=0, help="the run's first steps, left out of the timed set")
    ap.add_argument("--large-n", type=int, default=1 << 20)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the large terrain scene")
    ap.add_argument("--no-gas", action="store_true", help="skip the 1M sphere gas")
    ap.add_argument("--n", type=int, default=1 << 20, help="spheres in the gas")
    ap.add_argument("--gas-steps", type=int, default=30)
    ap.add_argument("--gas-warmup", type=int, default=5)
    ap.add_argument("--no-pile", action="store_true")
    ap.add_argument("--pile-n", type=int, default=65536)
    ap.add_argument("--pile-steps", type=int, default=600, help="timed steps of the pile's 600-step run")
    ap.add_argument("--no-queries", action="store_true", help="skip the scene queries on the settled pile")
    ap.add_argument("--no-config1", action="store_true", help="skip the 4096-sphere elastic gas")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--worlds", type=int, default=1024)
    ap.add_argument("--loop-worlds", type=int, default=128)
    ap.add_argument("--loop-steps", type=int, default=5)
    ap.add_argument("--batched-no-graph", action="store_true",
                    help="step the batch kernel by kernel after the autotuner instead of replaying HIP graphs "
                         "(for rocprofv3 kernel traces)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=65536, help="bodies in the CPU's contact-phase sample")
    ap.add_argument("--cpu-timeout", type=float, default=90.0)
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run one elementwise kernel of known traffic (2^25 doubles: reads 512 MiB, writes "
                         "256 MiB), the scale for rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("PROBE_FETCH", "PROBE_WRITE", "RUN_FETCH", "RUN_WRITE"),
                    help="print the HBM traffic summary (markdown) of four rocprofv3 --pmc counter_collection.csv "
                         "files: FETCH_SIZE and WRITE_SIZE of --traffic-probe, then of the profiled run")
    ap.add_argument("--kernel-summary", metavar="KERNEL_STATS_CSV",
                    help="print the per-kernel table (markdown) of a rocprofv3 --stats kernel_stats.csv")
    ap.add_argument("--cpu-sample", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-scene", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_sample:
        cpu_sample_main(args.cpu_sample, args.cpu_scene)
        return
    if args.pmc_summary:
        print(pmc_summary(*args.pmc_summary))
        return
    if args.kernel_summary:
        print(kernel_summary(args.kernel_summary))
        return
    if args.traffic_probe:
        import numpy as np
        from nativeengine_amd.gpu import GpuWorld
        rng = np.random.default_rng(0)
        a, b = rng.uniform(-1, 1, 1 << 25), rng.uniform(-1, 1, 1 << 25)
        out = GpuWorld().fp_probe(3, a, b)   # k_fp_probe: out = a*b + a
        assert np.array_equal(out, a * b + a)
        print(json.dumps({"traffic_probe": "k_fp_probe", "read_bytes": 2 * 8 << 25, "write_bytes": 8 << 25}))
        return

    if args.gpus != 1:
        print(json.dumps({"error": "the step runs on one GPU; sharding one world across GPUs is not "
                                   "implemented, so --gpus must be 1", "n_gpus": args.gpus}))
        sys.exit(2)

    metric = "body-steps/s at 1M bodies; batched world-steps/s; backward/forward time ratio"
    out = {"metric": metric, "value": None, "unit": "body-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic"}
    if not args.no_large:
        large = large_bench(args)
        log(f"large: {large['value']:.4g} body-steps/s, {large['ms_per_step']:.3f} ms/step")
        out.update({"value": large["value"], "ms_per_step": large["ms_per_step"], "steps": large["steps"],
                    "warmup": large["warmup"],
                    "config": {"workload": large["workload"], "bodies": large["bodies"], "parallelism": "single GPU"},
                    "roofline": large["roofline"], "cpu_baseline": large["cpu_baseline"],
                    "speedup_vs_cpu": large["speedup_vs_cpu"],
                    "large": {k: v for k, v in large.items() if k not in ("roofline", "cpu_baseline", "workload")}})
    else:
        out["note"] = "large terrain scene skipped (--no-large)"
    if not args.no_gas:
        gas = gas_bench(args)
        log(f"gas: {gas['value']:.4g} body-steps/s, {gas['ms_per_step']:.3f} ms/step")
        out["gas"] = gas
    if not args.no_pile:
        pile = pile_bench(args)
        log(f"pile: {pile['value']:.4g} body-steps/s, {pile['ms_per_step']:.3f} ms/step")
        out["pile"] = pile
    if not args.no_config1:
        c1 = config1_bench(args)
        log(f"gas4096: {c1['value']:.4g} body-steps/s, {c1['ms_per_step']:.3f} ms/step, dE/E {c1['energy_drift_rel']:.3e} "
            f"(reference {c1['reference_energy_drift_rel']:.3e}), bitwise {c1['bitwise_equal_to_reference']}")
        out["gas4096"] = c1
    if not args.no_batched:
        batched = batched_bench(args.worlds, args.loop_worlds, args.loop_steps, graph=not args.batched_no_graph)
        log(f"batched: {batched['value']:.4g} world-steps/s, {batched['ms_per_step']:.3f} ms/step, "
            f"{batched['vs_loop']:.1f}x the loop")
        out["batched"] = batched
    if not args.no_adjoint:
        adjoint = adjoint_bench(256, 120, 16)
        log(f"adjoint: forward {adjoint['forward_s']:.3f}s backward {adjoint['backward_s']:.3f}s "
            f"ratio {adjoint['value']:.2f}")
        out["adjoint"] = adjoint
    print(json.dumps(out))


if __name__ == "__main__":
    main()
