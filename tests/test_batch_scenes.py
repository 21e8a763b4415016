"""Batched-world scenes (CPU): concatenation and slicing are inverse, world
views carry their own box and material, and the oracle steps a world the
same whether it was built alone or cut out of a batch."""
import numpy as np

import engines
from nativeengine_amd import scene as S


def test_slice_of_concat_is_identity():
    a = S.mixed_pile(30, seed=1)
    b = S.mixed_pile(20, seed=2)
    c = S.concat_scenes([a, b])
    assert c.n == a.n + b.n
    for s, (lo, hi) in ((a, (0, a.n)), (b, (a.n, a.n + b.n))):
        t = S.slice_scene(c, lo, hi)
        for k in ("x", "v", "q", "w", "inv_mass", "inv_inertia", "shape", "radius", "half_extents", "flags",
                  "hull_offsets", "hull_verts", "nocollide_offsets", "nocollide"):
            assert np.array_equal(getattr(t, k), getattr(s, k)), k


def test_batched_worlds_layout():
    b = S.batched_worlds(12, seed=1)
    sizes = np.diff(b.offsets)
    assert b.offsets[0] == 0 and b.offsets[-1] == b.scene.n
    assert sizes.min() >= 64 and sizes.max() <= 4096
    assert np.all((b.friction >= 0) & (b.friction <= 0.8)) and np.all((b.restitution >= 0) & (b.restitution <= 1))
    w = b.world(4)
    assert w.params.half == tuple(b.half[4]) and w.params.periodic
    assert w.params.restitution == b.restitution[4] and w.params.friction == b.friction[4]
    # about half spheres, half boxes; 20% volume fraction
    n_sph = np.count_nonzero(w.shape == 0)
    assert abs(n_sph - w.n / 2) <= 1
    vol = np.sum(4.0 / 3.0 * np.pi * w.radius[w.shape == 0] ** 3) + np.sum(8 * np.prod(w.half_extents[w.shape == 2], 1))
    assert abs(vol / (2 * w.params.half[0]) ** 3 - 0.2) < 1e-9


def test_adjoint_worlds_layout():
    b = S.adjoint_worlds(3, n_spheres=10, n_boxes=6, side=16.0)
    assert list(np.diff(b.offsets)) == [17, 17, 17]
    g = b.scene
    assert np.all(g.inv_mass[b.offsets[:-1]] == 0.0)          # ground first, static
    assert np.all(g.half_extents[b.offsets[:-1]][:, 0] == 8.0)


def test_oracle_world_from_batch_equals_world_built_alone(oracle_lib):
    b = S.batched_worlds(4, seed=7, nmin=64, nmax=200)
    alone = b.world(2)
    cut = S.slice_scene(b.scene, int(b.offsets[2]), int(b.offsets[3]))
    cut.params = alone.params
    o1, o2 = engines.OracleWorld(), engines.OracleWorld()
    o1.set_scene(alone)
    o2.set_scene(cut)
    o1.step(5)
    o2.step(5)
    for u, v in zip(o1.state(), o2.state()):
        assert np.array_equal(u, v)
