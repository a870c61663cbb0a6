"""ctypes wrapper of the oracle library (oracle/lib/libnsl_oracle.so).

TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference step that the HIP
path is checked against, and the timed CPU baseline. Only tests/, the smoke check and
bench.py's CPU-baseline leg use it.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libnsl_oracle.so")
_LIB = None

BOX, SPHERE, CAPSULE = 0, 1, 2


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    return LIB_PATH


def lib():
    global _LIB
    if _LIB is None:
        build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        F, I = ctypes.c_float, ctypes.c_int
        fp = ctypes.POINTER(ctypes.c_float)
        L.nslo_create.restype = P
        L.nslo_destroy.argtypes = [P]
        L.nslo_set_options.argtypes = [P, I, I, I, I]
        L.nslo_set_gravity.argtypes = [P, F, F, F]
        L.nslo_set_restitution.argtypes = [P, I, F, I]
        L.nslo_create_body.argtypes = [P, I, ctypes.POINTER(ctypes.c_double), fp, fp, fp, I, F, F, F, I, F, F, I, I, I]
        lp = ctypes.POINTER(ctypes.c_int64)
        L.nslo_create_box.argtypes = [P, I, F, F, F, fp, lp]
        L.nslo_create_sphere.argtypes = [P, I, fp, F, fp, lp]
        L.nslo_create_capsule.argtypes = [P, I, fp, fp, F, fp, lp]
        L.nslo_step.argtypes = [P, F, I]
        L.nslo_set_residuals.argtypes = [P, I]
        L.nslo_get_residuals.argtypes = [P, ctypes.POINTER(ctypes.c_double), fp, fp, ctypes.POINTER(ctypes.c_int32)]
        L.nslo_state_hash.argtypes = [P]
        L.nslo_state_hash.restype = ctypes.c_uint64
        L.nslo_body_capacity.argtypes = [P]
        L.nslo_contact_count.argtypes = [P]
        L.nslo_get_bodies.argtypes = [P, ctypes.POINTER(ctypes.c_double), fp, fp, fp]
        L.nslo_get_stats.argtypes = [P, ctypes.POINTER(ctypes.c_int64)]
        L.nslo_contact_capacity.argtypes = [P]
        ip = ctypes.POINTER(ctypes.c_int32)
        L.nslo_get_contact_colors.argtypes = [P, ip, ip, ip]
        L.nslo_set_threads.argtypes = [ctypes.c_int]
        L.nslo_get_contact_shapes.argtypes = [P, ip, ip]
        L.nslo_get_shape_bounds.argtypes = [P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint8)]
        L.nslo_get_phase_ms.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
        L.nslo_max_threads.restype = ctypes.c_int
        L.nslo_create_revolute_joint.argtypes = [P, I, I, fp, fp, F, F, I, I, F, F, F, I, F, F, I, F, F]
        L.nslo_get_joint.argtypes = [P, I, fp]
        L.nslo_save_state.argtypes = [P]
        L.nslo_save_state.restype = P
        L.nslo_restore_state.argtypes = [P, P]
        L.nslo_time_of_impact.argtypes = [fp, I, F, fp, fp, I, F, fp, F, fp, ctypes.POINTER(ctypes.c_int32)]
        L.nslo_body_location.argtypes = [P, I, ctypes.POINTER(ctypes.c_int32)]
        L.nslo_island_summary.argtypes = [P, ctypes.POINTER(ctypes.c_int64)]
        L.nslo_awake_body_count.argtypes = [P]
        L.nslo_island_hash.argtypes = [P, I]
        L.nslo_island_hash.restype = ctypes.c_uint64
        L.nslo_first_island_contact.argtypes = [P, ctypes.POINTER(ctypes.c_int32)]
        L.nslo_unlink_and_split.argtypes = [P, I, I]
        L.nslo_check_islands.argtypes = [P]
        dp = ctypes.POINTER(ctypes.c_double)
        L.nslo_rollout_record.argtypes = [P, F, I, I]
        L.nslo_rollout_record.restype = P
        L.nslo_rollout_destroy.argtypes = [P]
        L.nslo_rollout_bodies.argtypes = [P]
        L.nslo_rollout_status.argtypes = [P]
        L.nslo_rollout_gradient.argtypes = [P, fp, dp, dp, dp, dp, ctypes.POINTER(ctypes.c_int)]
        L.nslo_rollout_loss.argtypes = [P, fp, I, ctypes.c_double, I]
        L.nslo_rollout_loss.restype = ctypes.c_double
        _LIB = L
    return _LIB


def set_threads(n: int) -> None:
    """Threads of the oracle's OpenMP loops (its results do not depend on this)."""
    lib().nslo_set_threads(int(n))


def max_threads() -> int:
    return int(lib().nslo_max_threads())


TOI_STATES = ("unknown", "failed", "overlapped", "hit", "separated")


def time_of_impact(points_a, radius_a, sweep_a, points_b, radius_b, sweep_b, max_fraction=1.0) -> dict:
    """time_of_impact (time_of_impact.rs:406-598) on two point proxies. A sweep is
    (local_center, c1, c2, q1, q2) with quaternions as (x, y, z, s)."""
    L = lib()
    fp = ctypes.POINTER(ctypes.c_float)
    pa = np.ascontiguousarray(points_a, np.float32).reshape(-1, 3)
    pb = np.ascontiguousarray(points_b, np.float32).reshape(-1, 3)
    sa = np.concatenate([np.asarray(x, np.float32).ravel() for x in sweep_a])
    sb = np.concatenate([np.asarray(x, np.float32).ravel() for x in sweep_b])
    assert sa.size == 17 and sb.size == 17 and 1 <= len(pa) <= 8 and 1 <= len(pb) <= 8
    fout = np.zeros(8, np.float32)
    iout = np.zeros(4, np.int32)
    L.nslo_time_of_impact(pa.ctypes.data_as(fp), len(pa), float(radius_a), sa.ctypes.data_as(fp),
                          pb.ctypes.data_as(fp), len(pb), float(radius_b), sb.ctypes.data_as(fp), float(max_fraction),
                          fout.ctypes.data_as(fp), iout.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return {"state": TOI_STATES[int(iout[0])], "point": fout[0:3].copy(), "normal": fout[3:6].copy(),
            "fraction": np.float32(fout[6]), "distance": np.float32(fout[7]), "distance_iterations": int(iout[1]),
            "push_back_iterations": int(iout[2]), "root_iterations": int(iout[3])}


class _OracleSnapshot:
    def __init__(self, L, h):
        self.L, self.h = L, h

    def __del__(self):
        try:
            if self.h:
                self.L.nslo_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _shape_options(kw: dict) -> dict:
    """ShapeDef fields beyond density, friction and restitution, with their defaults."""
    out = {"rolling_resistance": 0.0, "category_bits": 1, "mask_bits": 0xFFFF_FFFF_FFFF_FFFF, "group_index": 0,
           "is_sensor": False, "enable_speculative_contact": True, "enable_hit_events": False}
    unknown = set(kw) - set(out)
    if unknown:
        raise TypeError(f"unknown shape fields {sorted(unknown)}")
    out.update(kw)
    return out


class OracleWorld:
    """The oracle's World: create bodies/shapes from a scene, step, hash."""

    def __init__(self, enable_sleep=True, enable_continuous=True, enable_warm_starting=True, enable_speculative=True):
        self.L = lib()
        self.h = ctypes.c_void_p(self.L.nslo_create())
        self.L.nslo_set_options(self.h, int(enable_sleep), int(enable_continuous), int(enable_warm_starting), int(enable_speculative))

    def close(self):
        if self.h:
            self.L.nslo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def create_body(self, body_type, position, rotation=(0, 0, 0, 1), v=(0, 0, 0), w=(0, 0, 0), enable_sleep=True,
                    linear_damping=0.0, angular_damping=0.0, gravity_scale=1.0, is_bullet=False, safety_factor=0.5,
                    sleep_threshold=0.05, motion_locks=0, allow_fast_rotation=False, enable_contact_recycling=True) -> int:
        p = (ctypes.c_double * 3)(*position)
        q = (ctypes.c_float * 4)(*rotation)
        lv = (ctypes.c_float * 3)(*v)
        av = (ctypes.c_float * 3)(*w)
        return self.L.nslo_create_body(self.h, body_type, p, q, lv, av, int(enable_sleep), linear_damping, angular_damping,
                                       gravity_scale, int(is_bullet), float(safety_factor), float(sleep_threshold),
                                       int(motion_locks), int(allow_fast_rotation), int(enable_contact_recycling))

    @staticmethod
    def _shape_def(density, friction, restitution, rolling_resistance, category_bits, mask_bits, group_index, is_sensor,
                   enable_speculative_contact, enable_hit_events):
        mat = (ctypes.c_float * 4)(density, friction, restitution, rolling_resistance)
        def s64(u):  # the u64 filter bits as the int64 the C array holds
            u &= 0xFFFF_FFFF_FFFF_FFFF
            return u - (1 << 64) if u >= (1 << 63) else u
        opt = (ctypes.c_int64 * 6)(s64(category_bits), s64(mask_bits), group_index, int(is_sensor),
                                   int(enable_speculative_contact), int(enable_hit_events))
        return mat, opt

    def create_box(self, body, hx, hy, hz, density=1000.0, friction=0.6, restitution=0.0, **shape) -> int:
        mat, opt = self._shape_def(density, friction, restitution, **_shape_options(shape))
        return self.L.nslo_create_box(self.h, body, hx, hy, hz, mat, opt)

    def create_sphere(self, body, center, radius, density=1000.0, friction=0.6, restitution=0.0, **shape) -> int:
        mat, opt = self._shape_def(density, friction, restitution, **_shape_options(shape))
        return self.L.nslo_create_sphere(self.h, body, (ctypes.c_float * 3)(*center), radius, mat, opt)

    def create_capsule(self, body, c1, c2, radius, density=1000.0, friction=0.6, restitution=0.0, **shape) -> int:
        mat, opt = self._shape_def(density, friction, restitution, **_shape_options(shape))
        return self.L.nslo_create_capsule(self.h, body, (ctypes.c_float * 3)(*c1), (ctypes.c_float * 3)(*c2), radius, mat, opt)

    def create_revolute_joint(self, j) -> int:
        """j: natsel_amd.scenes.JointSpec."""
        fa = (ctypes.c_float * 7)(*j.frame_a)
        fb = (ctypes.c_float * 7)(*j.frame_b)
        return self.L.nslo_create_revolute_joint(self.h, j.body_a, j.body_b, fa, fb, j.constraint_hertz,
                                                 j.constraint_damping_ratio, int(j.collide_connected), int(j.enable_spring),
                                                 j.hertz, j.damping_ratio, j.target_angle, int(j.enable_limit), j.lower_angle,
                                                 j.upper_angle, int(j.enable_motor), j.max_motor_torque, j.motor_speed)

    def joint(self, joint_id: int):
        """(impulses f32[9]: linear xyz, perpendicular xy, spring, motor, lower, upper; graph colour)."""
        out = np.zeros(9, np.float32)
        color = self.L.nslo_get_joint(self.h, joint_id, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return out, int(color)

    def step(self, dt=1.0 / 60.0, substeps=4):
        self.L.nslo_step(self.h, dt, substeps)

    def world_state_hash(self) -> int:
        return self.L.nslo_state_hash(self.h)

    def save_state(self):
        """A copy of the world (freed with the returned object)."""
        return _OracleSnapshot(self.L, self.L.nslo_save_state(self.h))

    def restore_state(self, s) -> None:
        self.L.nslo_restore_state(self.h, s.h)

    def contact_count(self) -> int:
        return int(self.L.nslo_contact_count(self.h))

    def body_arrays(self):
        n = self.L.nslo_body_capacity(self.h)
        pos = np.zeros((n, 3), np.float64)
        rot = np.zeros((n, 4), np.float32)
        v = np.zeros((n, 3), np.float32)
        w = np.zeros((n, 3), np.float32)
        fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
        self.L.nslo_get_bodies(self.h, pos.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), fp(rot), fp(v), fp(w))
        return pos, rot, v, w

    def body_location(self, body: int) -> tuple:
        """(solver set, index in it, island id) of a body: set 0 static, 2 awake, 3+ sleeping."""
        out = (ctypes.c_int32 * 3)()
        self.L.nslo_body_location(self.h, body, out)
        return tuple(out)

    def awake_body_count(self) -> int:
        return int(self.L.nslo_awake_body_count(self.h))

    def island_summary(self) -> dict:
        out = (ctypes.c_int64 * 5)()
        self.L.nslo_island_summary(self.h, out)
        return dict(zip(["island_slots", "islands", "awake_islands", "sleeping_sets", "split_island"], list(out)))

    def island_hash(self, with_bodies: bool = False) -> int:
        return int(self.L.nslo_island_hash(self.h, int(with_bodies)))

    def first_island_contact(self) -> tuple:
        out = (ctypes.c_int32 * 2)()
        self.L.nslo_first_island_contact(self.h, out)
        return tuple(out)

    def unlink_and_split(self, contact: int, island: int) -> None:
        self.L.nslo_unlink_and_split(self.h, contact, island)

    def check_islands(self) -> int:
        """0 when every island, solver-set and encoding record agrees; else the failed check."""
        return int(self.L.nslo_check_islands(self.h))

    def stats(self) -> dict:
        out = (ctypes.c_int64 * 11)()
        self.L.nslo_get_stats(self.h, out)
        keys = ["pairs_created", "contacts_destroyed", "begin_touch", "end_touch", "recycled", "sat_calls", "sat_hits",
                "fast_bodies", "overflow_constraints", "toi_calls", "toi_hits"]
        return dict(zip(keys, list(out)))

    def phase_ms(self) -> dict:
        out = (ctypes.c_double * 5)()
        self.L.nslo_get_phase_ms(self.h, out)
        return dict(zip(["pairs", "collide", "prepare", "solve", "finalize"], list(out)))

    def shape_bounds(self, n_shapes: int):
        fat = np.zeros((n_shapes, 6), np.float32)
        moved = np.zeros(n_shapes, np.uint8)
        self.L.nslo_get_shape_bounds(self.h, fat.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                     moved.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        return fat, moved

    def contact_shapes(self):
        n = self.L.nslo_contact_capacity(self.h)
        a, b = np.zeros(n, np.int32), np.zeros(n, np.int32)
        ip = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))  # noqa: E731
        self.L.nslo_get_contact_shapes(self.h, ip(a), ip(b))
        return a, b

    def set_residuals(self, on: bool) -> None:
        self.L.nslo_set_residuals(self.h, 1 if on else 0)

    def colour_residuals(self) -> list:
        """The last step's residual sweep per colour 0..22 (World::colour_residuals)."""
        s = (ctypes.c_double * 23)()
        mn, mf = (ctypes.c_float * 23)(), (ctypes.c_float * 23)()
        k = (ctypes.c_int32 * 23)()
        self.L.nslo_get_residuals(self.h, s, mn, mf, k)
        return [{"sum_normal": s[i], "max_normal": mn[i], "max_friction": mf[i], "constraints": k[i]} for i in range(23)]

    def contact_colors(self):
        n = self.L.nslo_contact_capacity(self.h)
        color, local, alive = (np.zeros(n, np.int32) for _ in range(3))
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))  # noqa: E731
        self.L.nslo_get_contact_colors(self.h, ip(color), ip(local), ip(alive))
        return color, local, alive


def load_scene(scene) -> OracleWorld:
    """Build an OracleWorld from a natsel_amd.scenes.Scene."""
    w = OracleWorld(enable_sleep=scene.enable_sleep, enable_continuous=scene.enable_continuous)
    w.L.nslo_set_gravity(w.h, *(float(v) for v in scene.gravity))
    w.L.nslo_set_restitution(w.h, int(scene.restitution_iterations), float(scene.restitution_threshold),
                             int(scene.restitution_propagation))
    ids = [w.create_body(b.body_type, b.position, b.rotation, b.linear_velocity, b.angular_velocity, b.enable_sleep,
                         b.linear_damping, b.angular_damping, b.gravity_scale, b.is_bullet, b.safety_factor,
                         b.sleep_threshold, b.motion_locks, b.allow_fast_rotation, b.enable_contact_recycling)
           for b in scene.bodies]
    for s in scene.shapes:
        extra = {"rolling_resistance": s.rolling_resistance, "category_bits": s.category_bits, "mask_bits": s.mask_bits,
                 "group_index": s.group_index, "is_sensor": s.is_sensor,
                 "enable_speculative_contact": s.enable_speculative_contact, "enable_hit_events": s.enable_hit_events}
        if s.kind == BOX:
            w.create_box(ids[s.body], *s.params, s.density, s.friction, s.restitution, **extra)
        elif s.kind == SPHERE:
            w.create_sphere(ids[s.body], s.params[:3], s.params[3], s.density, s.friction, s.restitution, **extra)
        else:
            w.create_capsule(ids[s.body], s.params[:3], s.params[3:6], s.params[6], s.density, s.friction, s.restitution,
                             **extra)
    for j in scene.joints:
        w.create_revolute_joint(dataclasses.replace(j, body_a=ids[j.body_a], body_b=ids[j.body_b]))
    return w


class Rollout:
    """The oracle's reference for the differentiable rollout (ref_adjoint.cpp): the scene
    stepped `steps` steps on the oracle with its tape recorded, the frozen-active-set model
    run in float (which must repeat every recorded step bit for bit) and in double under
    reverse-mode differentiation. Parameters per awake body i: 7 i + {0..2} linear velocity,
    {3..5} angular velocity, 6 friction."""

    def __init__(self, scene, steps: int):
        w = load_scene(scene)
        self.L = w.L
        self.h = self.L.nslo_rollout_record(w.h, float(scene.dt), int(scene.substeps), int(steps))
        w.close()
        self.bodies = int(self.L.nslo_rollout_bodies(self.h))
        self.status = int(self.L.nslo_rollout_status(self.h))

    def close(self):
        if self.h:
            self.L.nslo_rollout_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def gradient(self, target) -> dict:
        """Loss |com - target|^2 and its gradient with respect to every awake body's initial
        linear and angular velocity and friction; `exact`: the float model repeated every
        recorded step bit for bit."""
        t = np.ascontiguousarray(target, np.float32)
        n = self.bodies
        gv, gw, gf = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
        loss, exact = ctypes.c_double(), ctypes.c_int()
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
        status = self.L.nslo_rollout_gradient(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(loss),
                                              dp(gv), dp(gw), dp(gf), ctypes.byref(exact))
        return {"loss": loss.value, "grad_v": gv, "grad_w": gw, "grad_friction": gf, "exact": bool(exact.value),
                "status": int(status)}

    def loss(self, target, param: int = -1, delta: float = 0.0, frozen: bool = True) -> float:
        """The model's loss in double precision with parameter `param` moved by `delta`."""
        t = np.ascontiguousarray(target, np.float32)
        return float(self.L.nslo_rollout_loss(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), int(param),
                                              float(delta), int(frozen)))


# --- narrow phase -------------------------------------------------------------------------
SHAPE_KINDS = {"box": 0, "sphere": 1, "capsule": 2}


def shape_box(hx, hy, hz, p=(0.0, 0.0, 0.0), q=(0.0, 0.0, 0.0, 1.0)):
    """A box hull (make_transformed_box_hull, hull.rs:754-852) placed by (p, q) in its shape frame."""
    return ("box", np.array([hx, hy, hz, *p, *q], np.float32))


def shape_sphere(center, radius):
    return ("sphere", np.array([*center, radius], np.float32))


def shape_capsule(c1, c2, radius):
    return ("capsule", np.array([*c1, *c2, radius], np.float32))


class NarrowCache:
    """A SATCache (box-box) or SimplexCache (box-round) carried between calls, as the
    reference's contact carries it. f = separation or metric; i = {type, index_a, index_b,
    hit} or {count, index_a[4], index_b[4]}."""

    def __init__(self):
        self.f = np.zeros(1, np.float32)
        self.i = np.zeros(9, np.int32)

    @property
    def sat(self) -> dict:
        return {"separation": np.float32(self.f[0]), "type": int(self.i[0]), "index_a": int(self.i[1]),
                "index_b": int(self.i[2]), "hit": int(self.i[3])}

    @property
    def simplex(self) -> dict:
        return {"metric": np.float32(self.f[0]), "count": int(self.i[0]), "index_a": self.i[1:5].tolist(),
                "index_b": self.i[5:9].tolist()}


# SeparatingFeature (collision_types.rs:97-121)
SAT_INVALID, SAT_BACKSIDE, SAT_FACE_A, SAT_FACE_B, SAT_EDGE = 0, 1, 2, 3, 4


class Narrow:
    """The oracle's narrow-phase routines (oracle/ref_narrow_capi.cpp) on plain arrays. `path`
    selects the library: the oracle itself by default, or a build with one planted mistake."""

    def __init__(self, path: str | None = None):
        if path is None:
            self.L = lib()
        else:
            self.L = ctypes.CDLL(path)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int32)
        self.L.nslo_collide_pair.argtypes = [ctypes.c_int, fp, ctypes.c_int, fp, fp, fp, ip, fp, ip]
        self.L.nslo_separating_axis.argtypes = [fp, fp, fp, ctypes.c_int, fp, ip]
        self.L.nslo_shape_distance.argtypes = [fp, ctypes.c_int, fp, ctypes.c_int, fp, fp, ip, fp, ip]
        self.L.nslo_box_hull.argtypes = [fp, fp, fp, ip, fp]
        self._fp, self._ip = fp, ip

    def cache(self) -> NarrowCache:
        return NarrowCache()

    def collide(self, a, b, xf=((0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0)), cache: NarrowCache | None = None) -> dict:
        """The register table's routine for (a, b) with B placed in A's frame by xf = (p, q).
        Returns normal, points [(point, separation, (owner1, index1, owner2, index2), feature id)]."""
        cache = cache or NarrowCache()
        t = np.array([*xf[0], *xf[1]], np.float32)
        of = np.zeros(19, np.float32)
        oi = np.zeros(21, np.int32)
        fp, ip = self._fp, self._ip
        rc = self.L.nslo_collide_pair(SHAPE_KINDS[a[0]], a[1].ctypes.data_as(fp), SHAPE_KINDS[b[0]], b[1].ctypes.data_as(fp),
                                      t.ctypes.data_as(fp), cache.f.ctypes.data_as(fp), cache.i.ctypes.data_as(ip),
                                      of.ctypes.data_as(fp), oi.ctypes.data_as(ip))
        if rc != 0:
            raise ValueError(f"no collide routine for ({a[0]}, {b[0]})")
        n = int(oi[0])
        pts = [(of[3 + 4 * k:6 + 4 * k].copy(), np.float32(of[6 + 4 * k]), tuple(int(x) for x in oi[1 + 4 * k:5 + 4 * k]),
                int(np.uint32(oi[17 + k]))) for k in range(n)]
        return {"normal": of[0:3].copy(), "points": pts, "cache": cache}

    def separating_axis(self, a, b, xf, early_return: bool = True) -> dict:
        """compute_separating_axis + get_best_axis for two box hulls."""
        t = np.array([*xf[0], *xf[1]], np.float32)
        of = np.zeros(7, np.float32)
        oi = np.zeros(4, np.int32)
        fp, ip = self._fp, self._ip
        self.L.nslo_separating_axis(a[1].ctypes.data_as(fp), b[1].ctypes.data_as(fp), t.ctypes.data_as(fp), int(early_return),
                                    of.ctypes.data_as(fp), oi.ctypes.data_as(ip))
        return {"normal": of[0:3].copy(), "separation": np.float32(of[3]), "face_a": np.float32(of[4]),
                "face_b": np.float32(of[5]), "edge": np.float32(of[6]), "kind": int(oi[0]), "index_a": int(oi[1]),
                "index_b": int(oi[2]), "separated": int(oi[3])}

    def distance(self, points_a, points_b, xf, cache: NarrowCache | None = None) -> dict:
        """shape_distance (GJK, use_radii false) between two point clouds, B placed by xf."""
        cache = cache or NarrowCache()
        pa = np.ascontiguousarray(points_a, np.float32).reshape(-1, 3)
        pb = np.ascontiguousarray(points_b, np.float32).reshape(-1, 3)
        t = np.array([*xf[0], *xf[1]], np.float32)
        of = np.zeros(10, np.float32)
        oi = np.zeros(1, np.int32)
        fp, ip = self._fp, self._ip
        self.L.nslo_shape_distance(pa.ctypes.data_as(fp), len(pa), pb.ctypes.data_as(fp), len(pb), t.ctypes.data_as(fp),
                                   cache.f.ctypes.data_as(fp), cache.i.ctypes.data_as(ip), of.ctypes.data_as(fp),
                                   oi.ctypes.data_as(ip))
        return {"point_a": of[0:3].copy(), "point_b": of[3:6].copy(), "normal": of[6:9].copy(), "distance": np.float32(of[9]),
                "iterations": int(oi[0]), "cache": cache}

    def box_hull(self, a) -> dict:
        pts = np.zeros((8, 3), np.float32)
        planes = np.zeros((6, 4), np.float32)
        edges = np.zeros((24, 4), np.int32)
        center = np.zeros(3, np.float32)
        fp, ip = self._fp, self._ip
        self.L.nslo_box_hull(a[1].ctypes.data_as(fp), pts.ctypes.data_as(fp), planes.ctypes.data_as(fp),
                             edges.ctypes.data_as(ip), center.ctypes.data_as(fp))
        return {"points": pts, "planes": planes, "edges": edges, "center": center}
