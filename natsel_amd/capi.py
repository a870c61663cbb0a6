"""ctypes binding of include/natsel_gpu.h (libnatsel_gpu.so, built in-tree)."""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def library_path() -> str:
    return os.path.join(_HERE, "lib", "libnatsel_gpu.so")


class Options(ctypes.Structure):
    _fields_ = [
        ("gravity", ctypes.c_float * 3),
        ("contact_hertz", ctypes.c_float),
        ("contact_damping_ratio", ctypes.c_float),
        ("contact_speed", ctypes.c_float),
        ("max_linear_speed", ctypes.c_float),
        ("restitution_threshold", ctypes.c_float),
        ("restitution_iterations", ctypes.c_int32),
        ("enable_sleep", ctypes.c_int32),
        ("enable_continuous", ctypes.c_int32),
        ("enable_warm_starting", ctypes.c_int32),
        ("enable_speculative", ctypes.c_int32),
        ("solver_blocks", ctypes.c_int32),
        ("enable_restitution_propagation", ctypes.c_int32),
        ("small_colour_items", ctypes.c_int32),
        ("stage_records", ctypes.c_int32),
        ("plain_step_launch", ctypes.c_int32),
    ]


class BodyDef(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32),
        ("position", ctypes.c_double * 3),
        ("rotation", ctypes.c_float * 4),
        ("linear_velocity", ctypes.c_float * 3),
        ("angular_velocity", ctypes.c_float * 3),
        ("linear_damping", ctypes.c_float),
        ("angular_damping", ctypes.c_float),
        ("gravity_scale", ctypes.c_float),
        ("enable_sleep", ctypes.c_int32),
        ("is_bullet", ctypes.c_int32),
        ("safety_factor", ctypes.c_float),
        ("sleep_threshold", ctypes.c_float),
        ("motion_locks", ctypes.c_int32),
        ("is_awake", ctypes.c_int32),
        ("is_enabled", ctypes.c_int32),
        ("allow_fast_rotation", ctypes.c_int32),
        ("enable_contact_recycling", ctypes.c_int32),
    ]


LOCK_LINEAR_X, LOCK_LINEAR_Y, LOCK_LINEAR_Z = 0x1, 0x2, 0x4  # MotionLocks bits (world_runtime.rs:87-92)
LOCK_ANGULAR_X, LOCK_ANGULAR_Y, LOCK_ANGULAR_Z = 0x8, 0x10, 0x20


class Material(ctypes.Structure):
    _fields_ = [
        ("density", ctypes.c_float),
        ("friction", ctypes.c_float),
        ("restitution", ctypes.c_float),
        ("rolling_resistance", ctypes.c_float),
        ("filter_category_bits", ctypes.c_uint64),
        ("filter_mask_bits", ctypes.c_uint64),
        ("filter_group_index", ctypes.c_int32),
        ("is_sensor", ctypes.c_int32),
        ("enable_speculative_contact", ctypes.c_int32),
        ("enable_hit_events", ctypes.c_int32),
    ]


class RevoluteJointDef(ctypes.Structure):
    _fields_ = [
        ("body_a", ctypes.c_int32),
        ("body_b", ctypes.c_int32),
        ("local_frame_a", ctypes.c_float * 7),
        ("local_frame_b", ctypes.c_float * 7),
        ("constraint_hertz", ctypes.c_float),
        ("constraint_damping_ratio", ctypes.c_float),
        ("collide_connected", ctypes.c_int32),
        ("enable_spring", ctypes.c_int32),
        ("hertz", ctypes.c_float),
        ("damping_ratio", ctypes.c_float),
        ("target_angle", ctypes.c_float),
        ("enable_limit", ctypes.c_int32),
        ("lower_angle", ctypes.c_float),
        ("upper_angle", ctypes.c_float),
        ("enable_motor", ctypes.c_int32),
        ("max_motor_torque", ctypes.c_float),
        ("motor_speed", ctypes.c_float),
    ]


class Profile(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("pairs_ms", "collide_ms", "state_pass_ms", "solve_ms", "finalize_ms", "step_ms")] + [
        (n, ctypes.c_int32)
        for n in ("contacts", "touching", "colors_used", "overflow", "fast_bodies", "sat_calls", "sat_hits", "recycled")
    ]


KERNEL_CLASSES = ["init", "collide", "solve", "continuous", "prepare", "finalize"]
PHASE_CLASSES = ["prepare", "integrate_velocities", "warm_start", "push", "integrate_positions", "relax", "restitution",
                 "overflow", "store_finalize"]
HOST_PHASES = ["pairs", "collide", "state_pass", "solve_finalize"]


class KernelStats(ctypes.Structure):
    _fields_ = [
        ("ms", ctypes.c_double * len(KERNEL_CLASSES)),
        ("launches", ctypes.c_int64 * len(KERNEL_CLASSES)),
        ("phase_ms", ctypes.c_double * len(PHASE_CLASSES)),
        ("phase_count", ctypes.c_int64 * len(PHASE_CLASSES)),
    ] + [
        (n, ctypes.c_int64)
        for n in ("steps", "bodies", "collided", "constraints", "points", "overflow_points", "recycled", "sat_calls",
                  "sat_hits", "fast_bodies", "barriers", "solver_blocks", "pair_tests")
    ] + [("host_ms", ctypes.c_double * len(HOST_PHASES)), ("joints", ctypes.c_int64), ("toi_calls", ctypes.c_int64),
         ("toi_hits", ctypes.c_int64)]


class ToiInput(ctypes.Structure):
    _fields_ = [
        ("points_a", (ctypes.c_float * 3) * 8),
        ("points_b", (ctypes.c_float * 3) * 8),
        ("count_a", ctypes.c_int32),
        ("count_b", ctypes.c_int32),
        ("radius_a", ctypes.c_float),
        ("radius_b", ctypes.c_float),
        ("sweep_a", ctypes.c_float * 17),
        ("sweep_b", ctypes.c_float * 17),
        ("max_fraction", ctypes.c_float),
    ]


class ToiOutput(ctypes.Structure):
    _fields_ = [
        ("state", ctypes.c_int32),
        ("point", ctypes.c_float * 3),
        ("normal", ctypes.c_float * 3),
        ("fraction", ctypes.c_float),
        ("distance", ctypes.c_float),
        ("distance_iterations", ctypes.c_int32),
        ("push_back_iterations", ctypes.c_int32),
        ("root_iterations", ctypes.c_int32),
    ]


SHAPE_CAPSULE, SHAPE_HULL, SHAPE_SPHERE = 0, 3, 5  # ShapeType (shape_def.rs:68-82)


class ShapeGeometry(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("params", ctypes.c_float * 7), ("frame", ctypes.c_float * 7)]


class CollideCache(ctypes.Structure):
    _fields_ = [("value", ctypes.c_float), ("kind", ctypes.c_int32), ("index_a", ctypes.c_uint8 * 4),
                ("index_b", ctypes.c_uint8 * 4), ("hit", ctypes.c_int32)]


class CollideQuery(ctypes.Structure):
    _fields_ = [("a", ShapeGeometry), ("b", ShapeGeometry), ("transform_b_to_a", ctypes.c_float * 7),
                ("cache", CollideCache)]


class LocalManifold(ctypes.Structure):
    _fields_ = [("normal", ctypes.c_float * 3), ("point_count", ctypes.c_int32), ("point", (ctypes.c_float * 3) * 4),
                ("separation", ctypes.c_float * 4), ("feature_id", ctypes.c_uint32 * 4), ("cache", CollideCache)]


class Tuning(ctypes.Structure):
    _fields_ = [("solver_blocks", ctypes.c_int32), ("small_colour_items", ctypes.c_int32), ("stage_records", ctypes.c_int32),
                ("step_kernel_us", ctypes.c_double), ("previous_us", ctypes.c_double), ("candidates", ctypes.c_int32)]


class ColourResidual(ctypes.Structure):
    _fields_ = [("sum_normal", ctypes.c_double), ("max_normal", ctypes.c_float), ("max_friction", ctypes.c_float),
                ("constraints", ctypes.c_int32), ("reserved", ctypes.c_int32)]


GRAPH_COLOURS = 23  # NSL_GRAPH_COLOURS: colours 0..22 (23 is the overflow colour)


class BatchTuning(ctypes.Structure):
    _fields_ = [("collide_blocks", ctypes.c_int32), ("use_graph", ctypes.c_int32), ("candidates", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("step_ms", ctypes.c_double), ("previous_ms", ctypes.c_double)]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("last_step_ms", ctypes.c_double),
        ("body_steps", ctypes.c_int64),
        ("bodies", ctypes.c_int64),
        ("contact_capacity", ctypes.c_int64),
        ("error_worlds", ctypes.c_int32),
        ("first_error", ctypes.c_int32),
        ("contacts_collided", ctypes.c_int64),
        ("constraints", ctypes.c_int64),
        ("points", ctypes.c_int64),
        ("overflow_constraints", ctypes.c_int64),
    ]


# Every entry point declared in include/natsel_gpu.h, with its argument types.
_P = ctypes.c_void_p
_I = ctypes.c_int32
_F = ctypes.c_float
SIGNATURES = {
    "nsl_last_error": (ctypes.c_char_p, []),
    "nsl_default_options": (None, [ctypes.POINTER(Options)]),
    "nsl_world_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Options), ctypes.POINTER(_P)]),
    "nsl_world_destroy": (ctypes.c_int, [_P]),
    "nsl_world_set_options": (ctypes.c_int, [_P, ctypes.POINTER(Options)]),
    "nsl_create_body": (ctypes.c_int, [_P, ctypes.POINTER(BodyDef), ctypes.POINTER(_I)]),
    "nsl_create_box": (ctypes.c_int, [_P, _I, ctypes.POINTER(_F), ctypes.POINTER(Material), ctypes.POINTER(_I)]),
    "nsl_create_sphere": (ctypes.c_int, [_P, _I, ctypes.POINTER(_F), _F, ctypes.POINTER(Material), ctypes.POINTER(_I)]),
    "nsl_create_capsule": (
        ctypes.c_int,
        [_P, _I, ctypes.POINTER(_F), ctypes.POINTER(_F), _F, ctypes.POINTER(Material), ctypes.POINTER(_I)],
    ),
    "nsl_default_revolute_joint_def": (None, [ctypes.POINTER(RevoluteJointDef)]),
    "nsl_default_body_def": (None, [ctypes.POINTER(BodyDef)]),
    "nsl_default_material": (None, [ctypes.POINTER(Material)]),
    "nsl_time_of_impact": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ToiInput), ctypes.POINTER(ToiOutput), _I]),
    "nsl_collide": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(CollideQuery), ctypes.POINTER(LocalManifold), _I]),
    "nsl_collide_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(CollideQuery), ctypes.POINTER(LocalManifold), _I]),
    "nsl_create_revolute_joint": (ctypes.c_int, [_P, ctypes.POINTER(RevoluteJointDef), ctypes.POINTER(_I)]),
    "nsl_world_get_joint": (ctypes.c_int, [_P, _I, ctypes.POINTER(_F), ctypes.POINTER(_I)]),
    "nsl_world_joint_count": (ctypes.c_int, [_P, ctypes.POINTER(_I)]),
    "nsl_world_contact_count": (ctypes.c_int, [_P, ctypes.POINTER(_I)]),
    "nsl_world_save_state": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "nsl_world_restore_state": (ctypes.c_int, [_P, _P]),
    "nsl_snapshot_destroy": (ctypes.c_int, [_P]),
    "nsl_world_step": (ctypes.c_int, [_P, _F, _I]),
    "nsl_world_step_n": (ctypes.c_int, [_P, _F, _I, _I]),
    "nsl_world_body_capacity": (ctypes.c_int, [_P, ctypes.POINTER(_I)]),
    "nsl_world_get_bodies": (
        ctypes.c_int,
        [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_F), ctypes.POINTER(_F), ctypes.POINTER(_F)],
    ),
    "nsl_world_state_hash": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "nsl_world_contact_capacity": (ctypes.c_int, [_P, ctypes.POINTER(_I)]),
    "nsl_world_get_contact_colors": (ctypes.c_int, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "nsl_world_get_contact_shapes": (ctypes.c_int, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "nsl_world_body_location": (ctypes.c_int, [_P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "nsl_world_island_summary": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "nsl_world_check_islands": (ctypes.c_int, [_P, ctypes.POINTER(_I)]),
    "nsl_world_get_profile": (ctypes.c_int, [_P, ctypes.POINTER(Profile)]),
    "nsl_world_set_kernel_timing": (ctypes.c_int, [_P, _I]),
    "nsl_world_autotune": (ctypes.c_int, [_P, _F, _I, _I, ctypes.POINTER(Tuning)]),
    "nsl_world_set_residuals": (ctypes.c_int, [_P, _I]),
    "nsl_world_get_residuals": (ctypes.c_int, [_P, ctypes.POINTER(ColourResidual), _I]),
    "nsl_world_reset_kernel_stats": (ctypes.c_int, [_P]),
    "nsl_world_get_kernel_stats": (ctypes.c_int, [_P, ctypes.POINTER(KernelStats)]),
    "nsl_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Options), ctypes.POINTER(_P)]),
    "nsl_batch_destroy": (ctypes.c_int, [_P]),
    "nsl_batch_add_world": (ctypes.c_int, [_P, ctypes.POINTER(_I)]),
    "nsl_batch_create_body": (ctypes.c_int, [_P, _I, ctypes.POINTER(BodyDef), ctypes.POINTER(_I)]),
    "nsl_batch_create_box": (ctypes.c_int, [_P, _I, _I, ctypes.POINTER(_F), ctypes.POINTER(Material), ctypes.POINTER(_I)]),
    "nsl_batch_create_sphere": (
        ctypes.c_int,
        [_P, _I, _I, ctypes.POINTER(_F), _F, ctypes.POINTER(Material), ctypes.POINTER(_I)],
    ),
    "nsl_batch_create_capsule": (
        ctypes.c_int,
        [_P, _I, _I, ctypes.POINTER(_F), ctypes.POINTER(_F), _F, ctypes.POINTER(Material), ctypes.POINTER(_I)],
    ),
    "nsl_batch_step": (ctypes.c_int, [_P, _F, _I, _I]),
    "nsl_batch_record": (ctypes.c_int, [_P, _F, _I, _I, _I]),
    "nsl_batch_adjoint": (ctypes.c_int, [_P, ctypes.POINTER(_F), ctypes.POINTER(_F), ctypes.POINTER(_F), ctypes.POINTER(_F),
                                         ctypes.POINTER(_F), ctypes.POINTER(_I)]),
    "nsl_batch_model_losses": (ctypes.c_int, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_double), _I,
                                              ctypes.POINTER(_F), _I, ctypes.POINTER(ctypes.c_double)]),
    "nsl_batch_rollout_times": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "nsl_batch_world_count": (ctypes.c_int, [_P, ctypes.POINTER(_I)]),
    "nsl_batch_body_count": (ctypes.c_int, [_P, _I, ctypes.POINTER(_I)]),
    "nsl_batch_state_hashes": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "nsl_batch_get_bodies": (
        ctypes.c_int,
        [_P, _I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_F), ctypes.POINTER(_F), ctypes.POINTER(_F)],
    ),
    "nsl_batch_contact_capacity": (ctypes.c_int, [_P, _I, ctypes.POINTER(_I)]),
    "nsl_batch_get_contact_colors": (ctypes.c_int, [_P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "nsl_batch_get_contact_shapes": (ctypes.c_int, [_P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "nsl_batch_get_shape_bounds": (ctypes.c_int, [_P, _I, ctypes.POINTER(_F), ctypes.POINTER(ctypes.c_uint8)]),
    "nsl_batch_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(BatchStats)]),
    "nsl_batch_set_tuning": (ctypes.c_int, [_P, ctypes.POINTER(BatchTuning)]),
    "nsl_batch_get_tuning": (ctypes.c_int, [_P, ctypes.POINTER(BatchTuning)]),
    "nsl_batch_autotune": (ctypes.c_int, [_P, _F, _I, _I, ctypes.POINTER(BatchTuning)]),
}


def load_library():
    """Load the in-tree HIP library. Raises if it has not been built: there is no
    other implementation to fall back to."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() (make -C natsel_amd/csrc)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


class NatselError(RuntimeError):
    pass


def check(status: int) -> None:
    if status != 0:
        msg = _LIB.nsl_last_error().decode() if _LIB is not None else "?"
        raise NatselError(f"natsel_gpu status {status}: {msg}")
