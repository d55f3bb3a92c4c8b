"""ctypes mirror of ``include/pdfc.h``.

This is the binding a Python caller of the C ABI needs (the same structs the
C++ host layer fills). Field order and types must match the header exactly.
"""
import ctypes as C

import numpy as np

PDFC_OK = 0
PDFC_ERR_INVALID = -1
PDFC_ERR_DEVICE = -2
PDFC_ERR_NUMERIC = -3
PDFC_ERR_CAPACITY = -4

OBSTACLE_SPHERE = 0
OBSTACLE_PLANE = 1
OBSTACLE_MESH = 2
OBSTACLE_BOX = 3
OBSTACLE_ROTATING_SPHERE = 4

FP64 = 0
FP32_REFINE64 = 1

WARM_PLAIN = 0
WARM_PREVIOUS = 1
WARM_LEARNED = 2
MLP_FEATURES = 12
MLP_OUTPUTS = 3
MLP_PARAMS = 10448
MLP_NORM = 27


class Obstacle(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("material", C.c_int32),
        ("point", C.c_double * 3),
        ("normal", C.c_double * 3),
        ("radius", C.c_double),
        ("tolerance", C.c_double),
        ("n_mesh_vertices", C.c_size_t),
        ("n_mesh_triangles", C.c_size_t),
        ("mesh_triangles", C.POINTER(C.c_int32)),
        ("n_frames", C.c_size_t),
        ("mesh_frames", C.POINTER(C.c_double)),
        ("frame_times", C.POINTER(C.c_double)),
        ("size", C.c_double * 3),
        ("rotation_axis", C.c_double * 3),
        ("angular_speed", C.c_double),
        ("rotation_start", C.c_double),
    ]


class Scene(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_size_t),
        ("rest_positions", C.POINTER(C.c_double)),
        ("positions", C.POINTER(C.c_double)),
        ("n_triangles", C.c_size_t),
        ("triangles", C.POINTER(C.c_int32)),
        ("area_density", C.c_double),
        ("bending", C.c_double),
        ("stretching", C.c_double),
        ("material", C.c_int32),
        ("n_obstacles", C.c_size_t),
        ("obstacles", C.POINTER(Obstacle)),
        ("n_materials", C.c_size_t),
        ("friction", C.POINTER(C.c_double)),
        ("n_attachments", C.c_size_t),
        ("attachment_vertices", C.POINTER(C.c_int32)),
        ("attachment_positions", C.POINTER(C.c_double)),
        ("attachment_weights", C.POINTER(C.c_double)),
    ]


class Params(C.Structure):
    _fields_ = [
        ("time_step", C.c_double),
        ("local_global_iterations", C.c_int32),
        ("self_collision", C.c_int32),
        ("self_collision_tolerance", C.c_double),
        ("gravity", C.c_double * 3),
        ("air_damping", C.c_double),
        ("precision", C.c_int32),
        ("refinement_sweeps", C.c_int32),
        ("warm_start", C.c_int32),
        ("record_iterations", C.c_int32),
        ("residual_tolerance", C.c_double),
        ("max_iterations", C.c_int32),
        ("use_graph", C.c_int32),
        ("warm_hold", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Contact(C.Structure):
    _fields_ = [
        ("vertex", C.c_int64),
        ("obstacle", C.c_int32),
        ("active", C.c_int32),
        ("normal", C.c_double * 3),
        ("point", C.c_double * 3),
        ("mu", C.c_double),
        ("r_local", C.c_double * 3),
        ("speed", C.c_double * 3),
    ]


class SelfContact(C.Structure):
    _fields_ = [
        ("vertex", C.c_int64),
        ("partner", C.c_int64),
        ("face", C.c_int64 * 3),
        ("level", C.c_int32),
        ("active", C.c_int32),
        ("normal", C.c_double * 3),
        ("mu", C.c_double),
        ("r_local", C.c_double * 3),
    ]


class Candidate(C.Structure):
    _fields_ = [
        ("vertex", C.c_int64),
        ("triangle", C.c_int64),
        ("barycentric", C.c_double * 3),
    ]


class Info(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_int64),
        ("n_triangles", C.c_int64),
        ("n_springs", C.c_int64),
        ("n_bending", C.c_int64),
        ("nnz_K", C.c_int64),
        ("nnz_L", C.c_int64),
        ("nnz_F", C.c_int64),
        ("n_supernodes", C.c_int64),
        ("n_levels", C.c_int64),
        ("n_contacts", C.c_int64),
        ("n_self_contacts", C.c_int64),
        ("n_self_candidates", C.c_int64),
        ("n_self_levels", C.c_int64),
        ("max_front", C.c_int64),
        ("factor_seconds", C.c_double),
        ("bytes_per_iteration", C.c_double),
        ("bytes_solve", C.c_double),
        ("bytes_local", C.c_double),
        ("bytes_forward", C.c_double),
        ("bytes_backward", C.c_double),
        ("n_below_rows", C.c_int64),
        ("bytes_top", C.c_double),
        ("top_m", C.c_int64),
        ("n_levels_swept", C.c_int64),
        ("waves_forward", C.c_int32 * 32),
        ("waves_backward", C.c_int32 * 32),
        ("autotune_ms", C.c_double),
        ("leaf_size", C.c_int32),
        ("top_chunks", C.c_int32),
        ("lhs_seconds", C.c_double),
        ("steps_done", C.c_int64),
        ("top_fold", C.c_int32),
        ("launches_forward", C.c_int32),
        ("launches_backward", C.c_int32),
        ("parts_forward", C.c_int32 * 32),
        ("parts_backward", C.c_int32 * 32),
        ("split_backward", C.c_int32 * 32),
        ("top_pregather", C.c_int32),
    ]


class PhaseStats(C.Structure):
    _fields_ = [
        ("step_us", C.c_double),
        ("detection_us", C.c_double),
        ("iterations_us", C.c_double),
        ("steps", C.c_int64),
        ("iterations", C.c_int64),
        ("iterations_to_tolerance", C.c_int64),
        ("contact_steps", C.c_int64),
        ("converged_steps", C.c_int64),
        ("fallbacks", C.c_int64),
        ("collision_us", C.c_double),
        ("timed_steps", C.c_int64),
        ("pairing_us", C.c_double),
        ("order_us", C.c_double),
        ("upload_us", C.c_double),
        ("contacts_sum", C.c_int64),
        ("self_contacts_sum", C.c_int64),
        ("self_levels_max", C.c_int64),
    ]


class KernelTimes(C.Structure):
    _fields_ = [
        ("local_us", C.c_double),
        ("forward_us", C.c_double),
        ("backward_us", C.c_double),
        ("top_us", C.c_double),
        ("forward_sweep_us", C.c_double),
        ("backward_sweep_us", C.c_double),
        ("local_launches", C.c_int64),
        ("forward_launches", C.c_int64),
        ("backward_launches", C.c_int64),
        ("top_launches", C.c_int64),
        ("self_us", C.c_double),
        ("self_iterations", C.c_int64),
    ]


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array (or NULL for None)."""
    if a is None:
        return C.POINTER(C.c_double)()
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def fptr(a):
    """Pointer to a C-contiguous float32 numpy array (or NULL for None)."""
    if a is None:
        return C.POINTER(C.c_float)()
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_float))
