"""Python mirror of the reference's World API over the HIP library.

Names follow the reference (src/simulation/world_runtime.rs, shape_creation.rs,
body_api.rs, solver_integration.rs) so code written against the CPU stepper reads the
same: World.create_body(BodyDef), create_hull_shape(body, ShapeDef, make_box_hull(...)),
create_sphere_shape, create_capsule_shape, step(time_step, sub_step_count),
world_state_hash(). Every call goes through include/natsel_gpu.h; nothing here
computes physics.
"""
from __future__ import annotations

import ctypes
import dataclasses
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .scenes import BOX, CAPSULE, DYNAMIC, KINEMATIC, SPHERE, STATIC, Scene  # noqa: F401


@dataclass
class BodyDef:  # src/types/body_def.rs:58-104 (fields and defaults)
    type: int = STATIC
    position: tuple = (0.0, 0.0, 0.0)
    rotation: tuple = (0.0, 0.0, 0.0, 1.0)
    linear_velocity: tuple = (0.0, 0.0, 0.0)
    angular_velocity: tuple = (0.0, 0.0, 0.0)
    linear_damping: float = 0.0
    angular_damping: float = 0.0
    gravity_scale: float = 1.0
    enable_sleep: bool = True
    is_bullet: bool = False
    safety_factor: float = 0.5
    sleep_threshold: float = 0.05
    motion_locks: int = 0  # capi.LOCK_* bits (MotionLocks)
    is_awake: bool = True
    is_enabled: bool = True
    allow_fast_rotation: bool = False
    enable_contact_recycling: bool = True


@dataclass
class ShapeDef:  # src/types/shape_def.rs:90-135 (fields and defaults; the base material's friction etc.)
    density: float = 1000.0
    friction: float = 0.6
    restitution: float = 0.0
    rolling_resistance: float = 0.0
    category_bits: int = 1  # Filter (shape_def.rs:18-24)
    mask_bits: int = 0xFFFF_FFFF_FFFF_FFFF
    group_index: int = 0
    is_sensor: bool = False
    enable_speculative_contact: bool = True
    enable_hit_events: bool = False


@dataclass
class BoxHull:
    """make_box_hull(hx, hy, hz) (src/collision/hull.rs:854)."""
    hx: float
    hy: float
    hz: float


def make_box_hull(hx: float, hy: float, hz: float) -> BoxHull:
    return BoxHull(hx, hy, hz)


@dataclass
class Sphere:
    center: tuple
    radius: float


@dataclass
class Capsule:
    center1: tuple
    center2: tuple
    radius: float


@dataclass
class WorldOptions:  # world_runtime.rs:1243-1304
    gravity: tuple = (0.0, -10.0, 0.0)
    contact_hertz: float = 30.0
    contact_damping_ratio: float = 10.0
    contact_speed: float = 3.0
    max_linear_speed: float = 400.0
    restitution_threshold: float = 1.0
    restitution_iterations: int = 2
    enable_sleep: bool = True
    enable_continuous: bool = True
    enable_warm_starting: bool = True
    enable_speculative: bool = True
    solver_blocks: int = 0  # workgroups of the step kernel; 0 = sized from the largest phase
    enable_restitution_propagation: bool = False
    small_colour_items: int = 128  # colours this small run on one workgroup (0: off); results do not change
    stage_records: bool = False  # the step kernel copies next-colour constraint records to LDS; results do not change
    plain_step_launch: bool = False  # an ordinary (not cooperative) launch of the same step-kernel grid


def _material(d: ShapeDef) -> capi.Material:
    return capi.Material(d.density, d.friction, d.restitution, d.rolling_resistance, d.category_bits, d.mask_bits,
                         d.group_index, int(d.is_sensor), int(d.enable_speculative_contact), int(d.enable_hit_events))


def c_body_def(d: BodyDef) -> capi.BodyDef:
    """BodyDef as the C ABI's nsl_body_def."""
    c = capi.BodyDef()
    c.type = d.type
    c.position[:] = d.position
    c.rotation[:] = d.rotation
    c.linear_velocity[:] = d.linear_velocity
    c.angular_velocity[:] = d.angular_velocity
    c.linear_damping, c.angular_damping, c.gravity_scale = d.linear_damping, d.angular_damping, d.gravity_scale
    c.enable_sleep = int(d.enable_sleep)
    c.is_bullet = int(d.is_bullet)
    c.safety_factor = d.safety_factor
    c.sleep_threshold = d.sleep_threshold
    c.motion_locks = int(d.motion_locks)
    c.is_awake, c.is_enabled = int(d.is_awake), int(d.is_enabled)
    c.allow_fast_rotation = int(d.allow_fast_rotation)
    c.enable_contact_recycling = int(d.enable_contact_recycling)
    return c


def c_options(o: WorldOptions) -> capi.Options:
    """WorldOptions as the C ABI's nsl_options."""
    c = capi.Options()
    c.gravity[:] = o.gravity
    c.contact_hertz, c.contact_damping_ratio, c.contact_speed = o.contact_hertz, o.contact_damping_ratio, o.contact_speed
    c.max_linear_speed, c.restitution_threshold = o.max_linear_speed, o.restitution_threshold
    c.restitution_iterations = o.restitution_iterations
    c.enable_sleep, c.enable_continuous = int(o.enable_sleep), int(o.enable_continuous)
    c.enable_warm_starting, c.enable_speculative = int(o.enable_warm_starting), int(o.enable_speculative)
    c.solver_blocks = int(o.solver_blocks)
    c.enable_restitution_propagation = int(o.enable_restitution_propagation)
    c.small_colour_items = int(o.small_colour_items)
    c.stage_records = int(o.stage_records)
    c.plain_step_launch = int(o.plain_step_launch)
    return c


class Snapshot:
    """A saved world state (device and host copies); freed with the object."""

    def __init__(self, lib, handle):
        self._lib, self._h = lib, handle

    def __del__(self):
        try:
            if self._h:
                self._lib.nsl_snapshot_destroy(self._h)
                self._h = None
        except Exception:
            pass


class World:
    def __init__(self, device: int = 0, options: WorldOptions | None = None):
        self._lib = capi.load_library()
        self._opts = options or WorldOptions()
        handle = ctypes.c_void_p()
        capi.check(self._lib.nsl_world_create(device, ctypes.byref(self._c_options()), ctypes.byref(handle)))
        self._h = handle

    def _c_options(self) -> capi.Options:
        return c_options(self._opts)

    @property
    def options(self) -> WorldOptions:
        return self._opts

    def apply_options(self) -> None:
        capi.check(self._lib.nsl_world_set_options(self._h, ctypes.byref(self._c_options())))

    def enable_sleeping(self, flag: bool) -> None:
        self._opts.enable_sleep = bool(flag)
        self.apply_options()

    def close(self) -> None:
        if self._h:
            capi.check(self._lib.nsl_world_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- creation -------------------------------------------------------------------
    def create_body(self, d: BodyDef) -> int:
        c = c_body_def(d)
        out = ctypes.c_int32()
        capi.check(self._lib.nsl_create_body(self._h, ctypes.byref(c), ctypes.byref(out)))
        return out.value

    def create_hull_shape(self, body: int, d: ShapeDef, hull: BoxHull) -> int:
        out = ctypes.c_int32()
        h = (ctypes.c_float * 3)(hull.hx, hull.hy, hull.hz)
        capi.check(self._lib.nsl_create_box(self._h, body, h, ctypes.byref(_material(d)), ctypes.byref(out)))
        return out.value

    def create_sphere_shape(self, body: int, d: ShapeDef, s: Sphere) -> int:
        out = ctypes.c_int32()
        c = (ctypes.c_float * 3)(*s.center)
        capi.check(self._lib.nsl_create_sphere(self._h, body, c, s.radius, ctypes.byref(_material(d)), ctypes.byref(out)))
        return out.value

    def create_capsule_shape(self, body: int, d: ShapeDef, c: Capsule) -> int:
        out = ctypes.c_int32()
        c1 = (ctypes.c_float * 3)(*c.center1)
        c2 = (ctypes.c_float * 3)(*c.center2)
        capi.check(self._lib.nsl_create_capsule(self._h, body, c1, c2, c.radius, ctypes.byref(_material(d)), ctypes.byref(out)))
        return out.value

    def create_revolute_joint(self, d) -> int:
        """World::create_revolute_joint (revolute_joint.rs:149-174). `d` has the fields of
        RevoluteJointDef (natsel_amd.scenes.JointSpec): body_a, body_b, frame_a/frame_b as
        (px, py, pz, qx, qy, qz, qs), constraint_hertz, constraint_damping_ratio,
        collide_connected, and the spring, limit and motor settings."""
        c = capi.RevoluteJointDef()
        c.body_a, c.body_b = int(d.body_a), int(d.body_b)
        c.local_frame_a[:] = d.frame_a
        c.local_frame_b[:] = d.frame_b
        c.constraint_hertz, c.constraint_damping_ratio = d.constraint_hertz, d.constraint_damping_ratio
        c.collide_connected = int(d.collide_connected)
        c.enable_spring, c.hertz, c.damping_ratio, c.target_angle = int(d.enable_spring), d.hertz, d.damping_ratio, d.target_angle
        c.enable_limit, c.lower_angle, c.upper_angle = int(d.enable_limit), d.lower_angle, d.upper_angle
        c.enable_motor, c.max_motor_torque, c.motor_speed = int(d.enable_motor), d.max_motor_torque, d.motor_speed
        out = ctypes.c_int32()
        capi.check(self._lib.nsl_create_revolute_joint(self._h, ctypes.byref(c), ctypes.byref(out)))
        return out.value

    def joint(self, joint_id: int):
        """(impulses f32[9]: linear xyz, perpendicular xy, spring, motor, lower, upper; graph colour)."""
        out = np.zeros(9, np.float32)
        color = ctypes.c_int32()
        capi.check(self._lib.nsl_world_get_joint(self._h, joint_id, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                 ctypes.byref(color)))
        return out, color.value

    def save_state(self) -> "Snapshot":
        """World::save_state (world_state.rs:340)."""
        h = ctypes.c_void_p()
        capi.check(self._lib.nsl_world_save_state(self._h, ctypes.byref(h)))
        return Snapshot(self._lib, h)

    def restore_state(self, s: "Snapshot") -> None:
        """World::restore_state (world_state.rs:439)."""
        capi.check(self._lib.nsl_world_restore_state(self._h, s._h))

    def contact_count(self) -> int:
        out = ctypes.c_int32()
        capi.check(self._lib.nsl_world_contact_count(self._h, ctypes.byref(out)))
        return out.value

    # --- stepping and queries ----------------------------------------------------------
    def step(self, time_step: float, sub_step_count: int) -> None:
        capi.check(self._lib.nsl_world_step(self._h, time_step, sub_step_count))

    def step_n(self, time_step: float, sub_step_count: int, steps: int) -> None:
        capi.check(self._lib.nsl_world_step_n(self._h, time_step, sub_step_count, steps))

    def world_state_hash(self) -> int:
        out = ctypes.c_uint64()
        capi.check(self._lib.nsl_world_state_hash(self._h, ctypes.byref(out)))
        return out.value

    def body_arrays(self):
        """(position f64 [n,3], rotation f32 [n,4], linear velocity [n,3], angular velocity [n,3])."""
        n = ctypes.c_int32()
        capi.check(self._lib.nsl_world_body_capacity(self._h, ctypes.byref(n)))
        n = n.value
        pos = np.zeros((n, 3), np.float64)
        rot = np.zeros((n, 4), np.float32)
        v = np.zeros((n, 3), np.float32)
        w = np.zeros((n, 3), np.float32)
        f = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))  # noqa: E731
        capi.check(self._lib.nsl_world_get_bodies(self._h, f(pos, ctypes.c_double), f(rot, ctypes.c_float), f(v, ctypes.c_float), f(w, ctypes.c_float)))
        return pos, rot, v, w

    def contact_colors(self):
        n = ctypes.c_int32()
        capi.check(self._lib.nsl_world_contact_capacity(self._h, ctypes.byref(n)))
        color, local, alive = (np.zeros(n.value, np.int32) for _ in range(3))
        f = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))  # noqa: E731
        capi.check(self._lib.nsl_world_get_contact_colors(self._h, f(color), f(local), f(alive)))
        return color, local, alive

    def contact_shapes(self):
        """Per contact id: the two shape ids (-1 when the id is free)."""
        n = ctypes.c_int32()
        capi.check(self._lib.nsl_world_contact_capacity(self._h, ctypes.byref(n)))
        a, b = np.zeros(n.value, np.int32), np.zeros(n.value, np.int32)
        f = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))  # noqa: E731
        capi.check(self._lib.nsl_world_get_contact_shapes(self._h, f(a), f(b)))
        return a, b

    def body_location(self, body: int) -> tuple:
        """(solver set, index in it, island id) of a body: set 0 static, 2 awake, 3+ sleeping."""
        out = [ctypes.c_int32() for _ in range(3)]
        capi.check(self._lib.nsl_world_body_location(self._h, int(body), *[ctypes.byref(x) for x in out]))
        return tuple(x.value for x in out)

    def body_is_awake(self, body: int) -> bool:
        """World::body_is_awake (body_api.rs:836-839)."""
        return self.body_location(body)[0] == 2

    def island_summary(self) -> dict:
        out = (ctypes.c_int64 * 5)()
        capi.check(self._lib.nsl_world_island_summary(self._h, out))
        return dict(zip(["island_slots", "islands", "awake_islands", "sleeping_sets", "split_island"], list(out)))

    def check_islands(self) -> int:
        """0 when the island and solver-set records agree; else the number of the failed check."""
        out = ctypes.c_int32()
        capi.check(self._lib.nsl_world_check_islands(self._h, ctypes.byref(out)))
        return out.value

    def set_kernel_timing(self, on: bool, phases: bool = True) -> None:
        """Launch events when `on`; the step kernel's per-phase clock as well when `phases`."""
        capi.check(self._lib.nsl_world_set_kernel_timing(self._h, (1 | (2 if phases else 0)) if on else 0))

    def autotune(self, time_step: float, sub_step_count: int, steps: int = 5) -> dict:
        """Times candidate launch shapes of the step kernel on `steps` steps from the current
        state (restored afterwards), keeps the fastest in the options and returns the choice;
        kernel statistics restart at zero."""
        t = capi.Tuning()
        capi.check(self._lib.nsl_world_autotune(self._h, time_step, sub_step_count, steps, ctypes.byref(t)))
        self._opts.solver_blocks = t.solver_blocks
        self._opts.small_colour_items = t.small_colour_items
        self._opts.stage_records = bool(t.stage_records)
        return {n: getattr(t, n) for n, _ in capi.Tuning._fields_}

    def set_residuals(self, on: bool) -> None:
        """Per-colour residual sweep after every step while on (a diagnostic: one more relax
        sweep over the coloured contacts on a copy of the body states; results unchanged)."""
        capi.check(self._lib.nsl_world_set_residuals(self._h, 1 if on else 0))

    def colour_residuals(self) -> list:
        """The last step's residuals per colour 0..22: dicts of sum_normal, max_normal,
        max_friction (impulse changes of one more relax sweep) and constraints."""
        arr = (capi.ColourResidual * capi.GRAPH_COLOURS)()
        capi.check(self._lib.nsl_world_get_residuals(self._h, arr, capi.GRAPH_COLOURS))
        return [{n: getattr(r, n) for n, _ in capi.ColourResidual._fields_ if n != "reserved"} for r in arr]

    def reset_kernel_stats(self) -> None:
        capi.check(self._lib.nsl_world_reset_kernel_stats(self._h))

    def kernel_stats(self) -> dict:
        k = capi.KernelStats()
        capi.check(self._lib.nsl_world_get_kernel_stats(self._h, ctypes.byref(k)))
        out = {n: getattr(k, n) for n, _ in capi.KernelStats._fields_ if n not in ("ms", "launches", "phase_ms", "phase_count", "host_ms")}
        out["ms"] = dict(zip(capi.KERNEL_CLASSES, list(k.ms)))
        out["launches"] = dict(zip(capi.KERNEL_CLASSES, list(k.launches)))
        out["phase_ms"] = dict(zip(capi.PHASE_CLASSES, list(k.phase_ms)))
        out["phase_count"] = dict(zip(capi.PHASE_CLASSES, list(k.phase_count)))
        out["host_ms"] = dict(zip(capi.HOST_PHASES, list(k.host_ms)))
        return out

    def profile(self) -> dict:
        p = capi.Profile()
        capi.check(self._lib.nsl_world_get_profile(self._h, ctypes.byref(p)))
        return {name: getattr(p, name) for name, _ in capi.Profile._fields_}


def body_def_of(b) -> BodyDef:
    """The BodyDef of a scene's BodySpec."""
    return BodyDef(b.body_type, b.position, b.rotation, b.linear_velocity, b.angular_velocity, b.linear_damping,
                   b.angular_damping, b.gravity_scale, b.enable_sleep, b.is_bullet, b.safety_factor,
                   sleep_threshold=b.sleep_threshold, motion_locks=b.motion_locks, allow_fast_rotation=b.allow_fast_rotation,
                   enable_contact_recycling=b.enable_contact_recycling)


def shape_def_of(s) -> ShapeDef:
    """The ShapeDef of a scene's ShapeSpec."""
    return ShapeDef(s.density, s.friction, s.restitution, s.rolling_resistance, s.category_bits, s.mask_bits, s.group_index,
                    s.is_sensor, s.enable_speculative_contact, s.enable_hit_events)


def load_scene(scene: Scene, device: int = 0, solver_blocks: int = 0, small_colour_items: int = 128,
               plain_step_launch: bool = False, stage_records: bool = False) -> World:
    """Build a World from a seeded scene description (natsel_amd.scenes)."""
    opts = WorldOptions(enable_sleep=scene.enable_sleep, enable_continuous=scene.enable_continuous, solver_blocks=solver_blocks,
                        small_colour_items=small_colour_items, plain_step_launch=plain_step_launch,
                        stage_records=stage_records,
                        gravity=tuple(scene.gravity), restitution_iterations=scene.restitution_iterations,
                        restitution_threshold=scene.restitution_threshold,
                        enable_restitution_propagation=scene.restitution_propagation)
    w = World(device, opts)
    ids = [w.create_body(body_def_of(b)) for b in scene.bodies]
    for s in scene.shapes:
        d = shape_def_of(s)
        if s.kind == BOX:
            w.create_hull_shape(ids[s.body], d, make_box_hull(*s.params))
        elif s.kind == SPHERE:
            w.create_sphere_shape(ids[s.body], d, Sphere(tuple(s.params[:3]), s.params[3]))
        else:
            w.create_capsule_shape(ids[s.body], d, Capsule(tuple(s.params[:3]), tuple(s.params[3:6]), s.params[6]))
    for j in scene.joints:
        w.create_revolute_joint(dataclasses.replace(j, body_a=ids[j.body_a], body_b=ids[j.body_b]))
    return w
