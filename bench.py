#!/usr/bin/env python3
"""bench.py -- body-steps/s of the MI355X step at 1M bodies; batched
world-steps/s; backward/forward time ratio of differentiable rollouts.

Headline (`value`): the large terrain scene, 1,048,576 bodies in a periodic
400 x 100 x 400 domain over a 1024 x 1024-vertex value-noise heightfield:
box columns of 8 (70%), loose spheres (25%), 16-link rope chains of rigid
distance joints (5%); friction 0.6, restitution 0.1, island sleeping on,
dt 1/60, 10 iterations, FP64, seed 5.  Inputs are synthetic and resident in
HBM before the timed region; a step is one full `World::step()`.  The timed
region is 300 single-step calls after 10 warm-up steps.  `cpu_baseline`
times the C restatement on one host core over the contact phase (one step at
steps 10/100/200/300 of the 65,536-body instance of the same scene, from the
GPU's state there); `cpu_start_phase` on the 1M scene's first two steps.

Other lines in the same JSON object:
  gas      the 1,048,576-sphere periodic elastic gas (r 0.5, L 253.98, e 1,
           mu 0, no gravity, 8 iterations, seed 0), the plain-method 1M point,
           with the reference itself (oracle/_ref) timed on the host;
  pile     the 65,536-body mixed-shape pile (40% boxes, 30% spheres, 20%
           capsules, 10% hulls) on a ground box inside 100 x 200 x 100 walls,
           10 iterations, 600 steps;
  batched  1024 periodic worlds of log-uniform size in [64, 4096] stepped in
           one set of launches (HIP-graph replay, autotuned) against looping
           the plain single-world step over a sample of the same worlds;
  adjoint  256 worlds of 256 spheres + 256 boxes on a ground box, 120 steps,
           checkpoint every 16: backward time / forward time.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-large] [--no-gas] [--no-pile]
                    [--no-batched] [--no-adjoint] [--traffic-probe]

Prints one JSON line.  `roofline` is measured live with HIP events on the
library's own stream around every launch of the dominant stage inside the
timed region, credited with the algorithmic bytes of every timed step;
`traffic` is the HBM bytes per launch of that stage from the committed
rocprofv3 PMC summary.
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


def scene_bench(s, steps: int, warmup: int, cpu_steps: int, cpu_what: str, label: str, max_seconds: float = 150.0,
                traffic_profile: str | None = None):
    """Body-steps/s of one resident scene over `steps` single-step calls after
    `warmup` untimed steps, with the dominant stage's HIP-event time and the
    algorithmic bytes it moved (units counted from every timed step)."""
    from nativeengine_amd.gpu import GpuWorld, KERNELS
    import numpy as np
    cpu = None
    if cpu_steps > 0:
        log(f"{label}: cpu baseline, {cpu_steps} steps ...")
        cpu = cpu_baseline(s, cpu_steps, cpu_what)
        if cpu:
            log(f"{label}: cpu {cpu['kind']} {cpu['value']:.4g} body-steps/s ({cpu['seconds']:.1f}s)")
    iters = int(s.params.solver_iterations)
    t0 = time.perf_counter()
    g = GpuWorld()
    g.set_scene(s)
    log(f"{label}: upload {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    for _ in range(warmup):
        g.step(1)
    log(f"{label}: warm-up {warmup} steps {time.perf_counter() - t0:.2f}s")
    stage_ms = {}
    for name in KERNELS:
        if name == "step":
            continue
        g.time_kernel(name)
        g.step(2)
        ms, k = g.kernel_time()
        stage_ms[name] = ms / max(k, 1)
    g.time_kernel(None)
    ranked = sorted(((v, k) for k, v in stage_ms.items() if stage_bytes(k, iters)), reverse=True)
    dom = ranked[0][1]
    log(f"{label}: stage ms/launch " + ", ".join(f"{k}={v:.4f}" for k, v in sorted(stage_ms.items(), key=lambda t: -t[1])))
    g.time_kernel(dom)
    units = {"body": 0, "pair": 0, "mesh_triangle": 0, "contact": 0}
    per_step_ms = []
    t0 = time.perf_counter()
    last_log = t0
    done = 0
    for _ in range(steps):
        ts = time.perf_counter()
        g.step(1)
        c = g.counters()
        now = time.perf_counter()
        per_step_ms.append((now - ts) * 1e3)
        done += 1
        units["body"] += s.n
        units["pair"] += c["candidate_pairs"]
        units["mesh_triangle"] += c["mesh_triangles"]
        units["contact"] += c["constraints"]
        if now - last_log > 10.0:
            log(f"{label}: {done} steps, {(now - t0) * 1e3 / done:.2f} ms/step, last {per_step_ms[-1]:.2f} ms, "
                f"{c['constraints']} constraints, largest island {c['max_island']}")
            last_log = now
        if now - t0 > max_seconds:
            log(f"{label}: stopping after {done} steps ({max_seconds:.0f} s cap)")
            break
    t1 = time.perf_counter()
    steps = done
    dom_ms, dom_launches = g.kernel_time()
    g.time_kernel(None)
    last = g.counters()
    reruns, trunc = g.epa_stats()
    stage_ms_end = {}   # the same stages at the end state (untimed steps after the timed region)
    for name in KERNELS:
        if name == "step":
            continue
        g.time_kernel(name)
        g.step(2)
        ms, k = g.kernel_time()
        stage_ms_end[name] = ms / max(k, 1)
    g.time_kernel(None)
    log(f"{label}: end-state stage ms/launch " +
        ", ".join(f"{k}={v:.3f}" for k, v in sorted(stage_ms_end.items(), key=lambda t: -t[1])))
    fl = g.body_aux()[0] if s.params.sleep_enabled else None
    g.close()
    value = s.n * steps / (t1 - t0)
    per = stage_bytes(dom, iters)
    bytes_total = sum(per[k] * units[k] for k in per)
    avg_launch_ms = dom_ms / max(dom_launches, 1)
    achieved = bytes_total / max(dom_launches, 1) / (avg_launch_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(dom, traffic_profile) if traffic_profile else None,
        "traffic_source": f"profiles/{traffic_profile} (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by a "
                          f"kernel of known traffic)" if traffic_profile else None,
        "bytes_per_unit": per,
        "units_per_launch": {k: round(units[k] / max(dom_launches, 1), 1) for k in per},
        "avg_launch_us": round(avg_launch_ms * 1e3, 3),
    }
    pct = np.percentile(per_step_ms, [50, 90, 100])
    out = {
        "value": value,
        "unit": "body-steps/s",
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": (t1 - t0) * 1e3 / steps,
        "step_ms_p50_p90_max": [round(float(x), 3) for x in pct],
        "bodies": s.n,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "stage_ms_end": {k: round(v, 4) for k, v in stage_ms_end.items()},
        "per_step_mean": {"candidate_pairs": units["pair"] / steps, "mesh_triangles": units["mesh_triangle"] / steps,
                          "constraints": units["contact"] / steps},
        "last_step": {"candidate_pairs": last["candidate_pairs"], "constraints": last["constraints"],
                      "islands": last["islands"], "largest_island": last["max_island"]},
        "epa_large_workspace_reruns": reruns,
        "epa_truncated": trunc,
    }
    if fl is not None:
        out["sleeping_bodies_at_end"] = int(np.sum((fl & 8) != 0))
    return out


def _save_state(g, s, path):
    """The world's bodies as the GPU holds them now (pose, velocity, sleep
    state) with the scene's shapes, materials, joints and meshes."""
    import numpy as np
    x, v, q, w = g.state()
    fl, tm, im, ii = g.body_aux()
    ims, iis = g.sleep_store()
    np.savez(path, x=x, v=v, q=q, w=w, flags=fl, timer=tm, im=im, ii=ii, ims=ims, iis=iis)


def _cpu_one_step(scene_args, state_path, timeout):
    """Seconds the C restatement takes for one step from a saved state, in a
    child process (a single step can hit the reference's non-converging EPA,
    which runs for minutes): None if it does not finish within `timeout`."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-sample", state_path, "--cpu-scene", scene_args]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return None
    if r.returncode != 0:
        log(r.stderr[-2000:])
        return None
    return float(r.stdout.strip().splitlines()[-1])


def cpu_sample_main(state_path, scene_args):
    """Child process: rebuild the scene, load the saved state, time one step."""
    import numpy as np
    from nativeengine_amd import scene as scn
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import engines
    n, seed = (int(t) for t in scene_args.split(","))
    s = scn.large_scene(n, seed=seed)
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


def contact_phase_cpu(n: int, marks, timeout: float):
    """The C restatement (single thread) on the n-body instance of the large
    scene at several points of its run, from the GPU's state there (the GPU's
    states are the restatement's, bit for bit): one step each."""
    import tempfile
    from nativeengine_amd import scene as scn
    from nativeengine_amd.gpu import GpuWorld
    s = scn.large_scene(n, seed=5)
    g = GpuWorld()
    g.set_scene(s)
    done, times = 0, []
    with tempfile.TemporaryDirectory() as tmp:
        for m in marks:
            g.step(m - done)
            done = m
            path = os.path.join(tmp, f"state_{m}.npz")
            _save_state(g, s, path)
            sec = _cpu_one_step(f"{n},5", path, timeout)
            log(f"contact-phase cpu: step {m}: " + (f"{sec:.2f}s" if sec is not None else f"> {timeout:.0f}s (stopped)"))
            times.append(sec)
    g.close()
    finished = [t for t in times if t is not None]
    # a step that did not finish counts as the timeout (the CPU figure is then an upper bound)
    total = sum(t if t is not None else timeout for t in times)
    return {
        "value": n * len(times) / total,
        "unit": "body-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"one step at each of steps {list(marks)} of the {n}-body instance of the same scene (same density, "
                  f"side {400.0 * (n / float(1 << 20)) ** 0.5:.0f}), each from the GPU's state at that step, single "
                  f"thread, {cpu_model()}, nproc={os.cpu_count()}",
        "seconds": total,
        "steps_over_timeout": len(times) - len(finished),
        "upper_bound": len(finished) < len(times),
    }


def large_bench(args):
    """The large terrain scene at 1,048,576 bodies: the headline line."""
    from nativeengine_amd import scene as scn
    t0 = time.perf_counter()
    s = scn.large_scene(args.large_n, seed=5)
    h = s.params.half
    log(f"large scene: {s.n} bodies, {len(s.joints)} joints, {len(s.meshes[0]['tris'])} triangles, "
        f"built in {time.perf_counter() - t0:.1f}s")
    r = scene_bench(s, args.large_steps, args.large_warmup, 0 if args.no_cpu else args.cpu_steps,
                    f"the same 1M-body scene from its initial state (bodies falling, no contacts yet: the CPU's "
                    f"fastest phase of the run)", "large", traffic_profile="large_terrain_1M_traffic.md")
    if not args.no_cpu:
        r["cpu_start_phase"] = r["cpu_baseline"]
        cp = contact_phase_cpu(args.cpu_n, [10, 100, 200, 300], args.cpu_timeout)
        r["cpu_contact_phases"] = cp
        r["cpu_baseline"] = cp
        r["speedup_vs_cpu"] = round(r["value"] / cp["value"], 1)
        log(f"large: cpu over the run {cp['value']:.4g} body-steps/s, GPU/CPU {r['speedup_vs_cpu']}")
    r["workload"] = (f"large terrain scene: {s.n} bodies (mesh + {int((s.shape == 2).sum())} boxes in columns of 8 + "
                     f"{int((s.shape == 0).sum())} spheres, {len(s.joints)} rope joints in 16-link ropes), periodic "
                     f"{2 * h[0]:.0f}x{2 * h[1]:.0f}x{2 * h[2]:.0f}, {len(s.meshes[0]['verts'])}-vertex value-noise "
                     f"heightfield amplitude 8, mu 0.6, e 0.1, sleeping on, dt 1/60, "
                     f"{s.params.solver_iterations} iterations, seed 5")
    return r


def pile_bench(args):
    """The mixed-shape pile under gravity (65,536 bodies, reflective walls)."""
    from nativeengine_amd import scene as scn
    t0 = time.perf_counter()
    s = scn.mixed_pile(args.pile_n, half=(50.0, 100.0, 50.0), seed=0, iterations=10)
    log(f"pile: {s.n} bodies, built in {time.perf_counter() - t0:.1f}s")
    r = scene_bench(s, args.pile_steps, args.pile_warmup, 0, "", "pile")
    r["workload"] = (f"mixed-shape pile: {s.n - 1} dynamic bodies (40% boxes, 30% spheres, 20% capsules, 10% "
                     f"12-vertex hulls) on a static ground box, reflective 100x200x100 box, g -9.81, mu 0.5, e 0.2, "
                     f"dt 1/60, 10 iterations, warm start, seed 0")
    return r


def batched_bench(worlds: int, steps: int, warmup: int, loop_worlds: int, loop_steps: int):
    """Batched world-steps/s against looping the plain step over `loop_worlds`
    of the same worlds (their own contexts, same state and warm cache)."""
    import numpy as np
    from nativeengine_amd import scene as scn
    from nativeengine_amd.gpu import GpuWorld
    t0 = time.perf_counter()
    b = scn.batched_worlds(worlds, seed=1)
    log(f"batched: {b.nworlds} worlds, {b.scene.n} bodies, built in {time.perf_counter() - t0:.1f}s")
    g = GpuWorld()
    g.set_scene(b.scene)
    g.set_worlds(b.offsets, b.half, b.restitution, b.friction)
    g.use_graph(True)
    t0 = time.perf_counter()
    for _ in range(warmup):
        g.step(1)
    log(f"batched warm-up {warmup} steps: {time.perf_counter() - t0:.1f}s")
    tune = g.autotune(steps=3)
    log(f"autotuner: {tune}")
    t0 = time.perf_counter()
    g.step(steps)
    t1 = time.perf_counter()
    reruns, trunc = g.epa_stats()
    batched_ms = (t1 - t0) * 1e3 / steps
    # the same state and warm cache, each sampled world in its own context
    x, v, q, w = g.state()
    keys, imp = g.export_warm_cache()
    ki, kj = (keys >> np.uint64(32)).astype(np.int64), (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    rng = np.random.default_rng(7)
    sample = np.sort(rng.choice(b.nworlds, size=min(loop_worlds, b.nworlds), replace=False))
    ctxs = []
    for wi in sample:
        a, e = int(b.offsets[wi]), int(b.offsets[wi + 1])
        ws = b.world(int(wi))
        ws.x, ws.v, ws.q, ws.w = x[a:e].copy(), v[a:e].copy(), q[a:e].copy(), w[a:e].copy()
        c = GpuWorld()
        c.set_scene(ws)
        sel = (ki >= a) & (ki < e)
        c.import_warm_cache(((ki[sel] - a).astype(np.uint64) << np.uint64(32)) | (kj[sel] - a).astype(np.uint64),
                            imp[sel])
        ctxs.append(c)
    t0 = time.perf_counter()
    for _ in range(loop_steps):
        for c in ctxs:
            c.step(1)
    t1 = time.perf_counter()
    loop_ms_per_world_step = (t1 - t0) * 1e3 / (loop_steps * len(ctxs))
    for c in ctxs:
        c.close()
    g.close()
    batched_wps = b.nworlds / (batched_ms * 1e-3)
    loop_wps = 1e3 / loop_ms_per_world_step
    return {
        "metric": "batched world-steps/s",
        "value": batched_wps,
        "unit": "world-steps/s",
        "ms_per_step": batched_ms,
        "steps": steps,
        "warmup": warmup,
        "worlds": b.nworlds,
        "bodies": int(b.scene.n),
        "body_steps_per_s": b.scene.n / (batched_ms * 1e-3),
        "loop_world_steps_per_s": loop_wps,
        "loop_sample": f"{len(ctxs)} of the worlds, each in its own context, {loop_steps} steps each, same state",
        "vs_loop": batched_wps / loop_wps,
        "autotune": tune,
        "epa_large_workspace_reruns": reruns,
        "epa_truncated": trunc,
        "config": "1024 periodic worlds, N_w log-uniform [64,4096] seed 1, half spheres r U[0.2,0.6] / half boxes "
                  "he U[0.2,0.6]^3, 20% volume fraction, mu_w U[0,0.8], e_w U[0,1], g=0, dt=1/60, 8 iterations",
    }


def adjoint_bench(worlds: int, steps: int, every: int):
    """Forward rollout and adjoint times on the differentiable-rollout workload."""
    import numpy as np
    from nativeengine_amd import scene as scn
    from nativeengine_amd.gpu import GpuWorld
    b = scn.adjoint_worlds(worlds, n_spheres=256, n_boxes=256, side=16.0, seed=2)
    nw = b.nworlds
    g = GpuWorld()
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
    ap.add_argument("--steps", type=int, default=300, help="timed steps of the large terrain scene")
    ap.add_argument("--warmup", type=int, default=10, help="untimed steps of the large terrain scene first")
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
    ap.add_argument("--pile-steps", type=int, default=600)
    ap.add_argument("--pile-warmup", type=int, default=0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--worlds", type=int, default=1024)
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--batched-warmup", type=int, default=30)
    ap.add_argument("--loop-worlds", type=int, default=128)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=65536, help="bodies in the CPU's contact-phase sample")
    ap.add_argument("--cpu-timeout", type=float, default=90.0)
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run one elementwise kernel of known traffic (2^25 doubles: reads 512 MiB, writes "
                         "256 MiB), the scale for rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--cpu-sample", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-scene", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_sample:
        cpu_sample_main(args.cpu_sample, args.cpu_scene)
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
    args.large_steps, args.large_warmup = args.steps, args.warmup

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
    if not args.no_batched:
        batched = batched_bench(args.worlds, args.batched_steps, args.batched_warmup, args.loop_worlds, 3)
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
