"""ctypes mirror of include/ne_gpu.h (parameter and body layouts, constants).

The same flat-array layout is used by the HIP library and by the test-side
engines, so one scene can be handed to each of them unchanged.
"""
import ctypes as C

import numpy as np

NE_OK = 0
NE_ERR_INVALID = -1
NE_ERR_HIP = -2
NE_ERR_CAPACITY = -3
NE_ERR_NONFINITE = -4
NE_ERR_UNSUPPORTED = -5
NE_ERR_NO_DEVICE = -6
NE_WARN_EPA_CAPPED = 1

SHAPE_SPHERE, SHAPE_CYLINDER, SHAPE_BOX, SHAPE_CAPSULE, SHAPE_CONVEX, SHAPE_PLANE, SHAPE_MESH, SHAPE_COMPOUND = range(8)

FLAG_KINEMATIC = 1
FLAG_SENSOR = 2
FLAG_DYNAMIC = 4
FLAG_SLEEPING = 8
FLAG_CCD = 16

# facade event kinds (PhysicsWorld::processContacts callbacks)
(NE_EVENT_CONTACT, NE_EVENT_BEGIN, NE_EVENT_STAY, NE_EVENT_END, NE_EVENT_TRIGGER_ENTER, NE_EVENT_TRIGGER_STAY,
 NE_EVENT_TRIGGER_EXIT) = range(7)

COMBINE_AVERAGE, COMBINE_MIN, COMBINE_MAX, COMBINE_MULTIPLY, COMBINE_GEOMETRIC_MEAN = range(5)

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)


class WorldParams(C.Structure):
    _fields_ = [
        ("half", C.c_double * 3),
        ("periodic", C.c_int32),
        ("open_boundary", C.c_int32),
        ("dt", C.c_double),
        ("gravity", C.c_double * 3),
        ("restitution", C.c_double),
        ("friction", C.c_double),
        ("contact_beta", C.c_double),
        ("restitution_combine", C.c_int32),
        ("friction_combine", C.c_int32),
        ("warm_start", C.c_int32),
        ("force_brute_force", C.c_int32),
        ("solver_iterations", C.c_int32),
        ("sleep_enabled", C.c_int32),
        ("sleep_lin_vel", C.c_double),
        ("sleep_ang_vel", C.c_double),
        ("sleep_time", C.c_double),
        ("wake_vel", C.c_double),
    ]


class Bodies(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("x", _dp), ("v", _dp), ("q", _dp), ("w", _dp),
        ("force", _dp), ("torque", _dp),
        ("inv_mass", _dp),
        ("inv_inertia", _dp),
        ("shape", _i32p),
        ("radius", _dp), ("half_height", _dp),
        ("half_extents", _dp),
        ("friction", _dp), ("restitution", _dp),
        ("linear_damping", _dp), ("angular_damping", _dp), ("gravity_scale", _dp),
        ("layer", _u32p), ("mask", _u32p),
        ("flags", _u32p),
        ("sleep_timer", _dp),
        ("inv_mass_store", _dp),
        ("inv_inertia_store", _dp),
        ("hull_offsets", _u32p),
        ("hull_verts", _dp),
        ("nocollide_offsets", _u32p),
        ("nocollide", _u32p),
    ]


class GpuConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("pair_capacity", C.c_uint32),
        ("contact_capacity", C.c_uint32),
        ("cell_scale", C.c_double),
        ("broadphase", C.c_int32),
        ("cells_per_body", C.c_uint32),
        ("epa_face_cap", C.c_uint32),
    ]

BROADPHASE_AUTO, BROADPHASE_WALK, BROADPHASE_BRICK, BROADPHASE_CELLS = 0, 1, 2, 3

# ne_distance_joint as a numpy record (80 bytes, the C layout)
JOINT_DTYPE = np.dtype([("a", "<u4"), ("b", "<u4"), ("local_a", "<f8", (3,)), ("local_b", "<f8", (3,)),
                        ("rest", "<f8"), ("stiffness", "<f8"), ("damping", "<f8")])


class Joint(C.Structure):
    """ne_joint: a general joint (World::joints)."""
    _fields_ = [
        ("type", C.c_int32), ("a", C.c_uint32), ("b", C.c_uint32),
        ("use_motor", C.c_int32), ("use_limit", C.c_int32), ("breakable", C.c_int32), ("broken", C.c_int32),
        ("local_a", C.c_double * 3), ("local_b", C.c_double * 3), ("axis_a", C.c_double * 3),
        ("axis_b", C.c_double * 3), ("ref_rel", C.c_double * 4),
        ("motor_speed", C.c_double), ("max_motor_impulse", C.c_double), ("lower", C.c_double), ("upper", C.c_double),
        ("break_impulse", C.c_double),
    ]


JOINT_BALL, JOINT_HINGE, JOINT_FIXED, JOINT_SLIDER = range(4)

# ne_joint as a numpy record (the C layout, offsets from the ctypes mirror)
_GJ = [("type", "<i4"), ("a", "<u4"), ("b", "<u4"), ("use_motor", "<i4"), ("use_limit", "<i4"), ("breakable", "<i4"),
       ("broken", "<i4"), ("local_a", ("<f8", (3,))), ("local_b", ("<f8", (3,))), ("axis_a", ("<f8", (3,))),
       ("axis_b", ("<f8", (3,))), ("ref_rel", ("<f8", (4,))), ("motor_speed", "<f8"), ("max_motor_impulse", "<f8"),
       ("lower", "<f8"), ("upper", "<f8"), ("break_impulse", "<f8")]
GJOINT_DTYPE = np.dtype({"names": [n for n, _ in _GJ], "formats": [f for _, f in _GJ],
                         "offsets": [getattr(Joint, n).offset for n, _ in _GJ], "itemsize": C.sizeof(Joint)})


class Mesh(C.Structure):
    """ne_mesh: a static triangle mesh attached to a NE_SHAPE_MESH body."""
    _fields_ = [
        ("body", C.c_uint32),
        ("nverts", C.c_uint32),
        ("verts", _dp),
        ("ntris", C.c_uint32),
        ("tris", _i32p),
    ]


class TuneResult(C.Structure):
    _fields_ = [
        ("cell_scale", C.c_double),
        ("broadphase", C.c_int32),
        ("graph", C.c_int32),
        ("ms_per_step", C.c_double),
        ("trials", C.c_uint32),
    ]


class QueryShape(C.Structure):
    """ne_query_shape: a query body for ne_gpu_overlap."""
    _fields_ = [
        ("shape", C.c_int32),
        ("nverts", C.c_uint32),
        ("x", C.c_double * 3),
        ("q", C.c_double * 4),
        ("half_extents", C.c_double * 3),
        ("radius", C.c_double),
        ("half_height", C.c_double),
        ("verts", C.POINTER(C.c_double)),
    ]


def dptr(a):
    """Pointer to a C-contiguous float64 array (or NULL for None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def uptr(a):
    if a is None:
        return None
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u32p)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i32p)
