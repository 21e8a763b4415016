"""Loads the golden fixtures written by tests/golden/make_fixtures.py."""
import glob
import json
import os

import numpy as np

from nativeengine_amd import scene as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def scene_names():
    return sorted(os.path.basename(p)[len("scene_"):-len(".npz")] for p in glob.glob(os.path.join(GOLDEN, "scene_*.npz")))


def _scene_from(z):
    n = len(z["in_inv_mass"])
    s = S.Scene(n)
    for k in z.files:
        if k.startswith("in_"):
            setattr(s, k[3:], z[k].copy())
    if "gjoints_raw" in z.files:
        from nativeengine_amd import _abi
        s.gjoints = z["gjoints_raw"].copy().view(_abi.GJOINT_DTYPE)
    p = json.loads(bytes(z["params_json"]).decode())
    p["half"] = tuple(p["half"])
    p["gravity"] = tuple(p["gravity"])
    s.params = S.Params(**p)
    return s


def load_scene(name):
    z = np.load(os.path.join(GOLDEN, f"scene_{name}.npz"), allow_pickle=False)
    s = _scene_from(z)
    out = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    if "diag_every" in z.files:
        out["diag_every"] = int(z["diag_every"][0])
    return s, int(z["steps"][0]), out


def load_snapshots():
    """The NES1 fixture: (scene, saves (k, bytes), save_every, load_blob,
    load_at, after_load (m, bytes), after_load_steps)."""
    z = np.load(os.path.join(GOLDEN, "nes1_sleeping_pile.npz"), allow_pickle=False)
    s = _scene_from(z)
    saves = [bytes(b) for b in z["saves"]]
    after = [bytes(b) for b in z["after_load"]]
    return (s, saves, int(z["save_every"][0]), bytes(z["load_blob"]), int(z["load_at"][0]), after,
            [int(k) for k in z["after_load_steps"]])


def load_narrow():
    z = np.load(os.path.join(GOLDEN, "narrow_pairs.npz"), allow_pickle=False)
    s = _scene_from(z)
    return s, z["pairs"], z["count"], z["offset"], z["normal"], z["point"], z["depth"]


def load_events(name):
    """A facade-event fixture: (scene, steps, events) with events[t][kind] a
    dict of i, j, point, normal, depth for step t (0-based)."""
    z = np.load(os.path.join(GOLDEN, f"events_{name}.npz"), allow_pickle=False)
    s = _scene_from(z)
    steps = int(z["steps"][0])
    off = z["ev_offset"]
    nk = (len(off) - 1) // steps
    ev = []
    for t in range(steps):
        row = []
        for k in range(nk):
            a, b = int(off[t * nk + k]), int(off[t * nk + k + 1])
            row.append({f: z[f"ev_{f}"][a:b] for f in ("i", "j", "point", "normal", "depth")})
        ev.append(row)
    out = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    return s, steps, ev, out


QUERY_NAMES = ["features", "pile_periodic", "special"]


def _shapes_from(z, prefix):
    from nativeengine_amd import queries as Q
    out = []
    voff = z[f"{prefix}_voff"]
    for k in range(len(z[f"{prefix}_shape"])):
        v = z[f"{prefix}_verts"][voff[k]:voff[k + 1]]
        out.append(Q.Shape(int(z[f"{prefix}_shape"][k]), tuple(z[f"{prefix}_x"][k]), q=tuple(z[f"{prefix}_q"][k]),
                           half_extents=tuple(z[f"{prefix}_he"][k]), radius=float(z[f"{prefix}_r"][k]),
                           half_height=float(z[f"{prefix}_hh"][k]), verts=v.copy() if len(v) else None))
    return out


def load_queries(name):
    """A scene-query fixture: (scene, state (x, v, q, w) the queries run on,
    inputs, expected).  inputs: org, dir, max_dist, mask, rad (rays; rad for
    the swept spheres), shapes0 / mask0 (overlap spheres and boxes), shapes1
    (overlapContacts bodies).  expected: ray / cast (nearest hits), all
    (raycastAll, CSR), ovl0 (overlaps, CSR), ovl1 (contacts, CSR)."""
    z = np.load(os.path.join(GOLDEN, f"queries_{name}.npz"), allow_pickle=False)
    s = _scene_from(z)
    state = tuple(z[f"state_{k}"] for k in "xvqw")
    inputs = dict(org=z["ray_org"], dir=z["ray_dir"], max_dist=z["ray_max"], mask=z["ray_mask"], rad=z["ray_rad"],
                  shapes0=_shapes_from(z, "ovl0"), mask0=z["ovl0_mask"], shapes1=_shapes_from(z, "ovl1"))
    hit = lambda p: dict(body=z[f"{p}_body"], distance=z[f"{p}_t"], point=z[f"{p}_p"], normal=z[f"{p}_n"])
    exp = dict(ray=hit("ray"), cast=hit("cast"),
               all=dict(offsets=z["all_off"], **hit("all")),
               ovl0=dict(offsets=z["ovl0_off"], body=z["ovl0_body"]),
               ovl1=dict(offsets=z["ovl1_off"], body=z["ovl1_body"], normal=z["ovl1_n"], depth=z["ovl1_d"]))
    return s, state, inputs, exp


QUERY_KINDS = ("ray", "cast", "all", "ovl0", "ovl1")


def run_queries(world, inputs, kinds=QUERY_KINDS):
    """The query kinds of a fixture's inputs on an engine with the query
    methods (GpuWorld, OracleWorld, RefWorld): results shaped as load_queries'
    expected."""
    i = inputs
    out = {}
    if "ray" in kinds:
        out["ray"] = world.raycast(i["org"], i["dir"], i["max_dist"], i["mask"])
    if "cast" in kinds:
        out["cast"] = world.sphere_cast(i["org"], i["rad"], i["dir"], i["max_dist"], i["mask"])
    if "all" in kinds:
        out["all"] = world.raycast_all(i["org"], i["dir"], i["max_dist"], i["mask"])
    if "ovl0" in kinds:
        o0 = world.overlap(i["shapes0"], i["mask0"])
        out["ovl0"] = dict(offsets=o0["offsets"], body=o0["body"])
    if "ovl1" in kinds:
        out["ovl1"] = world.overlap(i["shapes1"], None, contacts=True)
    return out


def same_queries(got, exp, what=""):
    """Bitwise equality of query results (every array of every kind)."""
    for kind, g in got.items():
        e = exp[kind]
        for f, a in e.items():
            b = g[f]
            b = np.ascontiguousarray(b).astype(a.dtype, copy=False)
            assert a.shape == b.shape, f"{what} {kind}.{f}: shape {b.shape} vs {a.shape}"
            if a.tobytes() != b.tobytes():
                rows = np.nonzero(np.any((a.view(np.uint8) != b.view(np.uint8)).reshape(len(a), -1), axis=1))[0]
                raise AssertionError(f"{what} {kind}.{f} differs in rows {rows[:8]} ({len(rows)} rows)")
