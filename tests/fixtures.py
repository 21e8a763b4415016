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
