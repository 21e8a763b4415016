"""The C++ drop-in (include/ne_world.hpp): compiles here against the header
and the HIP library; on the GPU runs restated selftest scenarios and a
compat-mode scene that must match the C restatement bit for bit."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import engines
from nativeengine_amd import scene as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "world_drop_in.cpp")
LIBDIR = os.path.join(ROOT, "nativeengine_amd")


def _build(tmp_path):
    exe = str(tmp_path / "world_drop_in")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
           SRC, "-o", exe, "-L", LIBDIR, "-lnegpu", f"-Wl,-rpath,{LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_drop_in_compiles(tmp_path):
    _build(tmp_path)


def test_adapter_compiles_against_reference_world(tmp_path):
    """The Session adapter takes the reference's own ne::World (src/world.hpp)
    unchanged: compile-only, where the reference sources are present."""
    ref = "/root/reference/src"
    if not os.path.exists(os.path.join(ref, "world.hpp")):
        pytest.skip("reference sources absent")
    prog = tmp_path / "adapt.cpp"
    prog.write_text('''
#include "world.hpp"
#include "ne_gpu_session.hpp"
void gpu_step(ne::World& w) { static thread_local ne_gpu::Session s; s.step(w, 1, 8); }
''')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", ref, "-I", os.path.join(ROOT, "include"), str(prog)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_drop_in_selftest_scenarios_on_gpu(tmp_path, gpu_available):
    exe = _build(tmp_path)
    r = subprocess.run([exe, "checks"], capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def _write_scene(path, s):
    p = s.params
    h = p.half
    hdr = [s.n, h[0], h[1], h[2], float(p.periodic), float(p.open_boundary), p.dt, *p.gravity, p.restitution,
           p.friction, p.contact_beta, float(p.restitution_combine), float(p.friction_combine), float(p.warm_start),
           float(p.solver_iterations), float(p.sleep_enabled), float(len(s.meshes))]
    hdr += [0.0] * (24 - len(hdr))
    with open(path, "wb") as f:
        f.write(struct.pack("24d", *hdr))
        for i in range(s.n):
            nv = int(s.hull_offsets[i + 1] - s.hull_offsets[i])
            row = [*s.x[i], *s.v[i], *s.q[i], *s.w[i], s.inv_mass[i], *s.inv_inertia[i], float(s.shape[i]),
                   s.radius[i], s.half_height[i], *s.half_extents[i], s.friction[i], s.restitution[i], float(nv)]
            f.write(struct.pack("26d", *row))
            for t in range(s.hull_offsets[i], s.hull_offsets[i + 1]):
                f.write(struct.pack("3d", *s.hull_verts[t]))
        for m in s.meshes:
            v, t = np.asarray(m["verts"], np.float64), np.asarray(m["tris"], np.float64)
            f.write(struct.pack("3d", float(m["body"]), float(len(v)), float(len(t))))
            f.write(np.ascontiguousarray(v).tobytes())
            f.write(np.ascontiguousarray(t).tobytes())


@pytest.mark.gpu
def test_drop_in_compat_mode_matches_oracle(tmp_path, gpu_available, oracle_lib):
    """World::step() in compat mode (upload, step, download every step) keeps
    the warm cache on the device and gives the oracle's bits."""
    exe = _build(tmp_path)
    s = S.mixed_pile(80, half=(4.0, 4.0, 4.0), seed=6)
    s.layer[:] = 0
    s.mask[:] = 0xFFFFFFFF
    s.flags[:] = np.where(s.inv_mass > 0, 4, 0).astype(np.uint32)   # dynamic / static
    k = 40
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_scene(fin, s)
    r = subprocess.run([exe, "scene", fin, fout, str(k)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(fout, np.float64).reshape(s.n, 16)[:, :13]
    o = engines.OracleWorld()
    o.set_scene(s)
    o.step(k)
    x, v, q, w = o.state()
    want = np.concatenate([x, v, q, w], axis=1)
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_drop_in_compat_mode_sleep_and_terrain(tmp_path, gpu_available, oracle_lib):
    """Compat mode over a terrain mesh with island sleeping on: the mesh goes
    up with the bodies, and the sleep state (flag, timer, zeroed and stored
    masses) comes back to the bodies and up again every step, so k single
    steps equal k resident steps of the oracle, sleep state included."""
    exe = _build(tmp_path)
    s = S.terrain_pile(30, seed=4, shapes="mixed")
    s.params.sleep_enabled = True
    k = 400                       # the oracle has 8 bodies asleep by then
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_scene(fin, s)
    r = subprocess.run([exe, "scene", fin, fout, str(k)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(fout, np.float64).reshape(s.n, 16)
    o = engines.OracleWorld()
    o.set_scene(s)
    o.step(k)
    x, v, q, w = o.state()
    assert np.array_equal(got[:, :13], np.concatenate([x, v, q, w], axis=1))
    fl, tm, im, _ = engines.body_aux(o, s.n)
    assert np.array_equal(got[:, 13], ((fl & 8) != 0).astype(np.float64))
    assert np.array_equal(got[:, 14], tm)
    assert np.array_equal(got[:, 15], im)
    assert np.any(got[:, 13] == 1.0), "some bodies fell asleep on the terrain"
