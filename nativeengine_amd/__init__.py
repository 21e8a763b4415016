"""nativeengine_amd -- MI355X-native (gfx950, HIP, FP64) per-step hot path of
NativeEngine: periodic minimum-image uniform-grid broadphase, narrow phase
(analytic sphere/cylinder/capsule, GJK/EPA, box-box clipping manifolds) and the
warm-started normal + Coulomb sequential-impulse solve with pose integration.

The product is the C-ABI library `libnegpu.so` (include/ne_gpu.h); this
package is its Python binding and scene tooling.
"""
from . import _abi  # noqa: F401
from .scene import (Body, Params, Scene, make_box, make_capsule, make_convex, make_cylinder,  # noqa: F401
                    make_plane, make_sphere, mixed_pile, sphere_gas)

__all__ = [
    "Body", "Params", "Scene", "make_box", "make_capsule", "make_convex", "make_cylinder", "make_plane",
    "make_sphere", "mixed_pile", "sphere_gas",
]
